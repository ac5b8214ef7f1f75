/*
 * kitamr_hip.h -- C ABI of the MI355X (gfx950) backend for KitAMR.jl's per-step
 * phase-space update: slope-limited reconstruction, the upwind DVM/CAIDVM face flux
 * (with interpolation across cells that carry different adapted velocity grids), the
 * Newton-Cotes moment sums, the Shakhov g+ build and the implicit relaxation update.
 *
 * The entry points replace the three per-step calls of the reference solver loop
 *     slope!(p4est,ka); flux!(p4est,ka); iterate!(p4est,ka)     src/Solver/Solver.jl:65-67
 * plus the data hand-over a Julia `ccall` shim performs after every
 * `amr_recover!` (src/Solver/AMR.jl:54-70) and before every host consumer of the state
 * (`check!`, `save_result`, AMR criteria).
 *
 * Conventions
 *   - Every function returns 0 on success and a negative KAH_E* code on failure; the
 *     message is kept in the context (kah_last_error). Nothing throws or aborts.
 *   - Host buffers belong to the caller and are copied in/out; calls return only when
 *     the caller's buffers may be reused. Device buffers belong to the library.
 *   - A context is bound to one thread and one device. There are no globals.
 *   - All indices are 0-based. Array layouts follow the reference's Julia column-major
 *     per-cell arrays, concatenated over cells in cell order:
 *        df   of cell c : df  [off_c*NDF     + k*n_c + i]          (VsData.df[n,NDF])
 *        sdf  of cell c : sdf [off_c*NDF*DIM + (d*NDF+k)*n_c + i]  (VsData.sdf[n,NDF,DIM])
 *        vs_mid cell c  : mid [off_c*DIM     + d*n_c + i]          (VsData.midpoint[n,DIM])
 *     with off_c = vs_offset[c], n_c = vs_offset[c+1]-off_c.
 */
#ifndef KITAMR_HIP_H
#define KITAMR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enumerations ------------------------------------------------------------- */
/* numerical flux, src/Abstract/Types.jl:85-91 (DVM, CAIDVM) */
enum { KAH_FLUX_CAIDVM = 0, KAH_FLUX_DVM = 1 };
/* time marching, src/Abstract/Types.jl:96-102 (CAIDVM_Marching, Euler, CIP_Marching; the last
 * is iterate!(::Type{CIP_Marching}), src/Theory/I-projection.jl:161-192) */
enum { KAH_MARCH_CAIDVM = 0, KAH_MARCH_EULER = 1, KAH_MARCH_CIP = 2 };
/* storage / arithmetic form of f on the device */
enum { KAH_PREC_FP64 = 0,    /* f, sdf, flux in FP64 (parity form)                      */
       KAH_PREC_FP32F = 1 }; /* f, sdf stored and advected in FP32, moments/w in FP64    */
/* domain boundary conditions, src/Abstract/Types.jl:125-153 */
enum { KAH_BC_MAXWELLIAN = 0, KAH_BC_SUPERSONIC_INFLOW = 1, KAH_BC_UNIFORM_OUTFLOW = 2,
       KAH_BC_INTERPOLATED_OUTFLOW = 3 };
/* neighbour states, src/Physical_space/Types.jl:1-22 / src/Mesh/Neighbor.jl:3-135:
 *   0 boundary, 1 same size, 2^(DIM-1) half size (finer), -1 double size (coarser) */

/* ---- error codes ---------------------------------------------------------------- */
enum { KAH_OK = 0, KAH_EINVAL = -1, KAH_ENOMEM = -2, KAH_EDEVICE = -3, KAH_ESTATE = -4,
       KAH_EUNSUPPORTED = -5 };

/* ---- parameters ------------------------------------------------------------------ */
typedef struct kah_params {
  int32_t dim;          /* 2 or 3                                                      */
  int32_t ndf;          /* 2 for DIM=2 (h,b: "2D2F"), 1 for DIM=3 ("3D1F")            */
  int32_t flux_kind;    /* KAH_FLUX_*                Solver.flux                       */
  int32_t marching;     /* KAH_MARCH_*               Solver.time_marching              */
  int32_t precision;    /* KAH_PREC_*                                                  */
  int32_t st_check_interval; /* Solver.ST_CHECK_INTERVAL (residual sampling)           */
  int32_t device;       /* HIP device ordinal                                          */
  int32_t reserved;
  double K;             /* Gas.K      internal degrees of freedom                      */
  double gamma;         /* Gas.γ                                                       */
  double Pr;            /* Gas.Pr                                                      */
  double mu_ref;        /* Gas.μᵣ as constructed by Julia (src/Gas/Types.jl:46)        */
  double omega;         /* Gas.ω                                                       */
} kah_params;

/* ---- flattened topology (one call after every adapt / recover) ------------------ */
typedef struct kah_topology {
  /* Cell records. Records [0, n_cells) are the local PsData cells (fluid: bound_enc>=0,
   * solid cells of an immersed boundary: bound_enc<0). Records [n_cells, n_cells+n_images)
   * are periodic images (src/Boundary/Period.jl:1-18): shallow copies of cell
   * image_src[k] that carry a shifted midpoint and share the source's velocity data,
   * flux and state. */
  int64_t n_cells;
  int64_t n_images;
  const int64_t* image_src;   /* [n_images]                                           */
  const double*  cell_mid;    /* [(n_cells+n_images)*dim]  PsData.midpoint             */
  const double*  cell_ds;     /* [(n_cells+n_images)*dim]  PsData.ds                   */
  const int32_t* bound_enc;   /* [n_cells]                 PsData.bound_enc            */
  const int32_t* ps_level;    /* [n_cells]  physical level, src/Flux/Slope.jl:825-829  */

  /* Velocity grids (VsData), concatenated in cell order. When vs_uniform is 1 every cell
   * carries the same grid (AMR_VS_MAXLEVEL == 0 builds identical grids for all cells,
   * src/Velocity_space/Initialize.jl:9-37): vs_level/vs_mid/vs_weight then hold that
   * one grid of n0 = vs_offset[1] cells and vs_offset[c] must equal c*n0. */
  const int64_t* vs_offset;   /* [n_cells+1]                                           */
  const int8_t*  vs_level;    /* [P]        VsData.level        ([n0] if vs_uniform)   */
  const double*  vs_mid;      /* [P*dim]    VsData.midpoint     ([n0*dim])             */
  const double*  vs_weight;   /* [P]        VsData.weight       ([n0])                 */
  int32_t vs_uniform;
  int32_t reserved0;

  /* Neighbours of each cell per face id (2*dim faces: xmin,xmax,ymin,ymax,zmin,zmax),
   * PsData.neighbor (src/Physical_space/Types.jl:15-22). nb_cell lists real cell ids
   * (fine neighbours in p4est sub-face order). Solid cells may leave this empty. */
  const int32_t* nb_state;    /* [n_cells*2*dim]                                       */
  const int64_t* nb_offset;   /* [n_cells*2*dim+1]                                     */
  const int64_t* nb_cell;     /* [nb_offset[n_cells*2*dim]]                            */

  /* Flux pairs: the reference's face list (src/Solver/Initialize.jl:5-228) with every
   * HangingFace/BackHangingFace expanded into its FluxData sub-faces, in face order.
   * here is always a local fluid cell record; there is a cell or image record, or -1
   * for a DomainFace. area is signed: rot*face_area (src/Flux/Flux.jl:64-111). */
  int64_t n_pairs;
  const int64_t* pair_here;   /* [n_pairs]                                             */
  const int64_t* pair_there;  /* [n_pairs]                                             */
  const int32_t* pair_dir;    /* [n_pairs]  0-based normal direction                   */
  const double*  pair_rot;    /* [n_pairs]  (-1)^(faceid-1) of the here side           */
  const double*  pair_area;   /* [n_pairs]                                             */
  const double*  pair_mid;    /* [n_pairs*dim] face (sub-face) midpoint                */
  const int32_t* pair_bc;     /* [n_pairs]  index into bc_* for domain faces, else -1  */

  /* Domain boundary conditions (Domain.bc, src/Boundary/Types.jl:12-33). */
  int32_t n_bc;
  int32_t vs_maxlevel;        /* Solver.AMR_VS_MAXLEVEL (key width)                    */
  const int32_t* bc_kind;     /* [n_bc]  KAH_BC_*                                      */
  const double*  bc_prim;     /* [n_bc*(dim+2)]  prim of the boundary state            */

  /* Velocity root grid (Configure.quadrature / vs_trees_num), used to key the cells of
   * adapted grids (src/Velocity_space/Neighbor.jl:9-24). */
  int32_t vs_trees_num[3];
  int32_t reserved;
  double  vs_lo[3];
  double  vs_hi[3];

  /* Immersed boundary (ImmersedBoundary, src/Boundary/Types.jl:185-199); n_ib_faces = 0
   * and n_ib_solid = 0 when there is none. The host computes the geometry once per mesh
   * (solid and donor flags in bound_enc, calc_intersect, the image stencils and the cut
   * velocity cells of initialize_cutted_velocity_cell); the library runs
   * update_solid_cell! and update_solid_neighbor! (src/Boundary/Immersed_boundary.jl:
   * 122-148, 436-500) on the device every step, between the slopes and the flux, on shared
   * or per-cell adapted velocity grids, with the CAIDVM or DVM flux; under CIP_Marching the
   * donors' update also runs positivity_preserving_ib! (src/Boundary/Positivity.jl).
   *
   * Solid cells: the bound_enc < 0 cells whose f update_solid_cell! rebuilds from their
   * fluid face neighbours (the first neighbour of each face id, in face-id order). */
  int64_t n_ib_solid;
  const int64_t* ib_solid;        /* [n_ib_solid]                                      */
  const int64_t* ib_solid_nb_off; /* [n_ib_solid+1]                                    */
  const int64_t* ib_solid_nb;     /* [ib_solid_nb_off[n_ib_solid]]                     */
  /* IB faces: one per (donor cell, solid face neighbour) = one SolidNeighbor. ib_face_pair
   * names the flux pair of the face (here = the donor, there = the solid cell). The image
   * stencil lists the donor's fluid neighbours (first per face id, face-id order) and then
   * the donor itself (image_df, :367-395). Cut cells: the velocity cells of the donor's
   * grid that u.n = 0 splits (indices into the donor's grid), with the areas (volumes in
   * 3D) of their gas (u.n < 0) and solid parts. */
  int64_t n_ib_faces;
  const int64_t* ib_face_pair;    /* [n_ib_faces]                                      */
  const double*  ib_aux_point;    /* [n_ib_faces*dim]  SolidNeighbor.aux_point          */
  const double*  ib_normal;       /* [n_ib_faces*dim]  SolidNeighbor.normal             */
  const double*  ib_bc;           /* [n_ib_faces*(dim+2)]  wall prim (get_bc); [0] unused */
  const int64_t* ib_image_off;    /* [n_ib_faces+1]                                    */
  const int64_t* ib_image_cell;   /* [ib_image_off[n_ib_faces]]                        */
  const int64_t* ib_cut_off;      /* [n_ib_faces+1]                                    */
  const int32_t* ib_cut_index;    /* [ib_cut_off[n_ib_faces]]  ascending per face       */
  const double*  ib_cut_gas;      /* CuttedVelocityCells.gas_weights                    */
  const double*  ib_cut_solid;    /* CuttedVelocityCells.solid_weights                  */
} kah_topology;

typedef struct kah_ctx kah_ctx;

/* ---- lifecycle -------------------------------------------------------------------- */
/* Replaces the KA/KInfo set-up of initialize(config), src/Solver/Initialize.jl:554-573 */
int  kah_create(const kah_params* params, kah_ctx** out);
/* Destroys a context and frees its device memory; accepts NULL. */
void kah_destroy(kah_ctx* ctx);
/* Message of the last failing call on ctx (never NULL). */
const char* kah_last_error(const kah_ctx* ctx);

/* Replaces the flattening point after initialize_faces!/update_faces!
 * (src/Solver/Initialize.jl:5-228, src/Solver/AMR.jl:54-70). Drops any previous state. */
int kah_set_topology(kah_ctx* ctx, const kah_topology* topo);

/* Uploads the cell state: w and prim [n_cells*(dim+2)], df [P*ndf] (layout above).
 * Entries of solid cells are copied as given (they are read as neighbours only).
 * Zeroes the flux accumulators (PsData.flux, VsData.flux). */
int kah_set_state(kah_ctx* ctx, const double* w, const double* prim, const double* df);

/* Downloads any of w, prim, qf [n_cells*(dim+2) | n_cells*dim], sw [n_cells*(dim+2)*dim]
 * (PsData.sw, the macro slope of src/Flux/Slope.jl:1022-1036), df [P*ndf] and
 * sdf [P*ndf*dim]. Pass NULL to skip an array. */
int kah_get_state(kah_ctx* ctx, double* w, double* prim, double* qf, double* sw,
                  double* df, double* sdf);

/* Downloads the flux accumulators PsData.flux [n_cells*(dim+2)] and VsData.flux [P*ndf]
 * (NULL skips). They hold the face fluxes between kah_flux and kah_iterate and are zero
 * after kah_iterate / kah_step, as after the reference's iterate!. */
int kah_get_flux(kah_ctx* ctx, double* flux, double* vs_flux);

/* Downloads the distribution of every SolidNeighbor, as update_solid_neighbor!
 * (src/Boundary/Immersed_boundary.jl:436-483) left it in the last flux: the wall-side data of
 * the IB faces, face after face, each [ndf][n_d] on its donor's grid (n_d velocity cells;
 * n0 each on a shared grid). KAH_ESTATE before the first flux or without immersed
 * boundary. */
int kah_get_ib(kah_ctx* ctx, double* aux_df);

/* Ghost-exchange buffers of a list of cells, in the reference's wire format, one buffer per
 * listed cell concatenated in list order (a cell may be listed more than once):
 *   KAH_MIRROR_DATA : [ds(dim) midpoint(dim) w(dim+2) vs_num bound_enc df(vs_num*ndf)]
 *                     (size_Ghost_Data; solid cells: w = Inf, vs_num = 0, no df)
 *   KAH_MIRROR_SLOPE: [sdf(vs_num*ndf*dim) sw((dim+2)*dim)], arrays column-major
 *                     (size_Ghost_Slope; solid cells: [EPS 0 ...])
 * replacing get_mirror_data / get_mirror_slope and update_mirror_data! /
 * update_mirror_slope! (src/Parallel/Ghost.jl:525-579, :757-808) on the sending side and the
 * ghost wrap of the receiving side (src/Parallel/Ghost.jl:688-749). */
enum { KAH_MIRROR_DATA = 0, KAH_MIRROR_SLOPE = 1 };
/* offsets[0..n]: start of each cell's buffer in doubles; offsets[n] is the total length. */
int kah_mirror_layout(kah_ctx* ctx, const int32_t* cells, int64_t n, int kind,
                      int64_t* offsets);
/* Packs the buffers of the listed cells from the device state into out (host memory,
 * offsets[n] doubles). SLOPE buffers carry sw recomputed from the current sdf. */
int kah_pack_mirrors(kah_ctx* ctx, const int32_t* cells, int64_t n, int kind, double* out);
/* Writes w and df (DATA) or sdf and sw (SLOPE) of the listed fluid cells from in (host memory,
 * the layout above); solid cells are skipped. DATA buffers are checked first: a solid buffer
 * (w = Inf) for a fluid cell, the converse, or a vs_num other than the cell's velocity-cell
 * count is KAH_EINVAL and nothing is written. FP32 contexts round df and sdf to float. */
int kah_unpack_mirrors(kah_ctx* ctx, const int32_t* cells, int64_t n, int kind,
                       const double* in);

/* ---- velocity-space adaptation on the device ------------------------------------- */
/* vs_adaptive_mesh_refinement! (src/Velocity_space/AMR.jl:168-181) with the distribution
 * kept in HBM: the criteria run on the device and only one flag byte per velocity cell
 * goes to the host, which rebuilds the grids (index work) and hands them back; df is
 * remapped and projected on the device. Parameters as in Solver (src/Solver/Types.jl:79-100). */
enum { KAH_VS_LOHNER = 0,          /* ADAPT_VS_MODE = :lohner (the default)            */
       KAH_VS_CONTRIBUTION = 1 };  /* the contribution criteria                         */
enum { KAH_VS_REFINE = 1, KAH_VS_COARSEN = 2 };   /* flag bits */
typedef struct kah_vs_adapt {
  int32_t mode;          /* KAH_VS_*                                                    */
  int32_t reserved;
  double coeff_local;    /* ADAPT_COEFFI_VS_LOCAL  (1e-2)                               */
  double coeff_global;   /* ADAPT_COEFFI_VS_GLOBAL (0.125)                              */
  double coeff_lohner;   /* ADAPT_COEFFI_VS_LOHNER (0.6)                                */
} kah_vs_adapt;
/* vr[0..1] = Velocity_Resolution (density, energy) of the current state, the maxima over
 * the fluid cells of vs_resolution (src/Velocity_space/AMR.jl:133-167). */
int kah_vs_resolution(kah_ctx* ctx, double* vr);
/* kah_vs_resolution for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): the call allocates nothing. No counterpart
 * in the reference. Its scratch (the per-cell partial maxima) is a fixed device array of
 * 16 MiB that each GPU the library runs on holds until the process ends, so it takes contexts
 * of at most 1,048,576 cells (KAH_EUNSUPPORTED beyond). Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (from two threads or two contexts) race. */
int kah_vs_resolution_noalloc(kah_ctx* ctx, double* vr);
/* flags[P]: per velocity cell, KAH_VS_REFINE as vs_refine! decides it (level below
 * vs_maxlevel and the criterion, src/Velocity_space/AMR.jl:24-65) and KAH_VS_COARSEN as
 * vs_coarsen!'s per-cell willingness (:69-111), both on the current grids, state, slopes,
 * w and prim, with the given vr (vs_adaptive_mesh_refinement! computes vr once, before
 * refining, and uses it for both passes). */
int kah_vs_flags(kah_ctx* ctx, const kah_vs_adapt* params, const double* vr, uint8_t* flags);
/* Replaces the velocity grids (new_topo: the same cells, faces and neighbours, new vs_*
 * arrays) and carries the state over on the device: for new velocity cell i of cell c,
 * src[off_new[c] + i] = j >= 0 takes the old cell j of c (kept cell, or a refined parent
 * injected into its children: refine_grid_stream!, src/Velocity_space/Rebuild.jl:44-85),
 * src = -(j+1) takes the mean of the 2^dim old cells j.. (coarsen_grid_stream!, :96-163).
 * Slopes of cells with changed[c] != 0 are zero afterwards (the rebuilt VsData), others
 * keep theirs; w and prim are kept; flux accumulators are zero. The sources are checked
 * before anything changes (KAH_EINVAL); a failure after that (device memory) leaves the
 * context without a state, to be set again. */
int kah_vs_remap(kah_ctx* ctx, const kah_topology* new_topo, const int32_t* src,
                 const uint8_t* changed);
/* conserved_I_porjection! of the listed cells' h onto their w (vs_conserved_correction!,
 * src/Velocity_space/AMR.jl:117-131; src/Theory/I-projection.jl:144-158); cells inside a
 * solid are skipped. */
int kah_vs_project(kah_ctx* ctx, const int32_t* cells, int64_t n);

/* Per-rank phase-space arrays of a restart checkpoint, as save_for_restart collects them
 * (src/IO/Restart.jl:80-117) and _restart_build_ps_list reads them (:306-321): N = n_cells,
 * Nv = sum of vs_num over the cells outside solids; matrices column-major:
 *   vs_nums [N] (0 inside a solid), bound_encs [N], ws [N x (dim+2)] (0 inside a solid),
 *   vs_levels [Nv], vs_midpoints [Nv x dim], vs_df [Nv x ndf].
 * The container (JLD2 files, the p4est forest, configuration and status) stays with the host. */
int kah_restart_sizes(kah_ctx* ctx, int64_t* n_cells, int64_t* n_vs);
int kah_save_restart(kah_ctx* ctx, int64_t* vs_nums, int64_t* bound_encs, double* ws,
                     int8_t* vs_levels, double* vs_midpoints, double* vs_df);
/* Sets the state from restart arrays of the same topology (vs_nums and bound_encs must match
 * it, else KAH_EINVAL and nothing is written): w = ws, prim = get_prim(w), df = vs_df, slopes
 * and flux accumulators zero. May replace kah_set_state. */
int kah_load_restart(kah_ctx* ctx, const int64_t* vs_nums, const int64_t* bound_encs,
                     const double* ws, const double* vs_df);

/* ---- the per-step path --------------------------------------------------------------- */
/* slope!(p4est,ka), src/Flux/Slope.jl:1047-1070: level-ordered velocity-space slopes
 * with the transverse correction on coarse-neighbour faces. dt is status.Δt of the step
 * (limit_Δt!, src/Theory/Iterate.jl:27-39); it is needed here because the per-face
 * macroscopic half fluxes are formed together with the slopes. */
int kah_slope(kah_ctx* ctx, double dt);
/* flux!(p4est,ka), src/Flux/Flux.jl:458-487 (non-IB faces + domain faces). Completes the
 * face fluxes into PsData.flux / VsData.flux. Must follow kah_slope with the same dt. */
int kah_flux(kah_ctx* ctx, double dt);
/* iterate!(p4est,ka), src/Theory/Iterate.jl:5-16 with the selected marching scheme.
 * step is status.step before the increment; when step % ST_CHECK_INTERVAL == 0 the
 * residual sums of src/Solver/Finalize.jl:5-11 are returned in sum_res/sum_avg
 * [dim+2] (may be NULL); otherwise they are zeroed. */
int kah_iterate(kah_ctx* ctx, double dt, int64_t step, double* sum_res, double* sum_avg);
/* slope + flux + iterate of one step in one call (fused flux/iterate kernels). Returns
 * without waiting for the device unless the residual is sampled this step; host buffers
 * are not touched before that point. */
int kah_step(kah_ctx* ctx, double dt, int64_t step, double* sum_res, double* sum_avg);

/* Waits for all work queued by kah_step on the context's stream. */
int kah_sync(kah_ctx* ctx);

/* ---- introspection -------------------------------------------------------------------- */
/* Number of phase-space updates per step: sum of vs_num over fluid cells. */
int64_t kah_phase_cells(const kah_ctx* ctx);
/* Device bytes held by the context (state + topology). */
int64_t kah_device_bytes(const kah_ctx* ctx);
/* Launch options (autotuned schedule choices). KAH_OPT_FLUX_WAVES: waves per workgroup of
 * the flux/relaxation kernel (2, 3, 4 or 8; 0, the default, chooses per size-class launch: for
 * the 2D FP32 form on adapted velocity grids 2 when the LDS copy of f* takes at most 8 KiB
 * (grids of up to 640 cells), 4 up to 40 KiB, else 8; 8 for 3D and the other forms);
 * a change drops captured graphs. KAH_OPT_SLOPE_VPT: velocity cells per thread of the slope
 * kernel (1..6, default 3 in 2D and 6 in 3D; its workgroups have 3 or 4 waves, per cell
 * whichever pads the grid least). KAH_OPT_VEL_TILES: 1 (default) processes a shared tensor-product velocity grid in
 * 8x8 / 4x4x4 tiles per wave, 0 in storage order. KAH_OPT_OVERLAP_MAPS: 1 (default)
 * precomputes, per distinct pair of adapted velocity grids that face each other, the
 * overlapping velocity cells of the other grid (4 bytes per velocity cell of the first), 0
 * searches the neighbour's keys in every step. These three take effect at the next
 * kah_set_topology. KAH_OPT_GRAPHS: 1 replays the launches of kah_step as one captured
 * HIP graph per buffer parity and dt (recaptured when dt changes; not while kernel timing is
 * on), 0 (default) launches them one by one. KAH_OPT_REGULAR: 1 (default) gives regular cells
 * a slope and a flux gather specialised to them, bit-identical to the general ones: for the
 * slope, cells with one neighbour per side and no transverse projection; for the gather
 * (CAIDVM flux), cells whose 2*DIM faces go to fluid cells on the same level, which read the
 * outgoing micro fluxes the slope kernel stores instead of reconstructing; 0 runs every cell
 * through the general slope and gather. Takes effect at the next kah_set_topology.
 * KAH_OPT_BRANCHES: 1..4 (default 4) streams over which the independent launches of a phase
 * (slope launches of one projection depth, flux/relaxation launches) are spread, forked from
 * and joined back to the context's stream by events (parallel branches of a captured graph).
 * No counterpart in the reference. */
enum { KAH_OPT_FLUX_WAVES = 1, KAH_OPT_SLOPE_VPT = 2, KAH_OPT_VEL_TILES = 3,
       KAH_OPT_OVERLAP_MAPS = 4, KAH_OPT_GRAPHS = 5, KAH_OPT_REGULAR = 6,
       KAH_OPT_BRANCHES = 7 };
int kah_set_option(kah_ctx* ctx, int option, int64_t value);
/* Times reps device-to-device copies of `bytes` bytes with a plain streaming copy kernel
 * (attainable HBM bandwidth = 2*bytes / ms_per_copy); the reference point for the
 * hardware byte counters. No counterpart in the reference (measurement only). */
int kah_copy_bandwidth(kah_ctx* ctx, int64_t bytes, int reps, double* ms_per_copy);
/* Average duration in ms of each kernel class over the calls since the last reset,
 * measured with HIP events on the library's stream: out[0] slope, out[1] flux+iterate. */
int kah_kernel_times(kah_ctx* ctx, int reset, double* out_ms, int64_t* out_calls);

#ifdef __cplusplus
}
#endif
#endif /* KITAMR_HIP_H */

// kah_api.cpp -- the C ABI of include/kitamr_hip.h: context, topology derivation, state
// transfer and the per-step launches.
//
// kah_set_topology turns the reference-shaped topology (neighbour lists with p4est states,
// the face list as flux pairs) into the per-cell descriptors the kernels gather from:
//   - a slope descriptor per (cell, direction), which resolves the reference's dispatch on
//     the neighbour-state pair (src/Flux/Slope.jl:458-771) and, for cells finer than the
//     coarsest level, the transverse-projection cases (src/Flux/Slope.jl:849-945);
//   - the half fluxes each cell computes (the moments of its own upwind micro flux through a
//     face, src/Flux/CAIDVM.jl:119) and the face sides each cell gathers in K2, in
//     face-list order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <limits>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/kitamr_hip.h"
#include "kah_internal.h"
#include "kah_kernels.h"

using namespace kah;

namespace {

struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

struct LevelBin {
  int level;
  int maxh;
  int wpi = 4;    // chunked kernel: waves per workgroup (3 or 4)
  int n;          // work items: cells (per-cell kernel) or (cell, chunk) pairs
  bool chunked;
  DBuf items;
};

struct EvPair {
  hipEvent_t a, b;
  int cls;
};

}  // namespace

struct kah_ctx {
  kah_params P{};
  hipStream_t stream = nullptr;
  std::string err;
  bool have_topo = false, have_state = false;
  int D = 2, NDF = 2, NW = 4;
  int64_t nc = 0, np = 0, npairs = 0;
  bool uni = true;
  bool f32 = false;   // FP32 form: f, sdf stored as float and advected in FP32
  int n0 = 0;
  int lmax = 0;
  int max_n = 0;
  size_t lds_bytes = 0;
  size_t lds_limit = 56 * 1024;   // dynamic LDS per workgroup (set from the device)
  int64_t phase_cells = 0;
  std::vector<int32_t> fluid;
  std::vector<int32_t> benc;
  std::vector<LevelBin*> bins;
  DBuf d_fluid, d_voff, d_gu, d_gw, d_glev, d_gkey, d_gof;
  int64_t n_grids = 1, grid_cells = 0, map_entries = 0;   // distinct grids, their cells, map entries
  DBuf d_df[2], d_sdf, d_oflux, d_vflux, d_psflux, d_w, d_prim, d_qf, d_half, d_rho, d_res, d_scratch;
  DBuf d_cpre;   // per-cell prologue records of the fused step (k_cell_pre)
  DBuf d_sl, d_slnb, d_hoff, d_halves, d_eoff, d_ents, d_recds, d_bc, d_sw, d_poff, d_pcnt;
  DBuf d_vtb, d_ovmap, d_moff;
  // ghost-exchange buffers: cell midpoints and bound_enc on the device, and the staging of
  // kah_pack_mirrors / kah_unpack_mirrors (grown on demand)
  std::vector<int64_t> voff_h;
  DBuf d_cmid, d_benc, d_xcells, d_xoff, d_xbuf, d_xfluid;
  // immersed boundary: solid cells with their fluid neighbours, the IB faces (their
  // SolidNeighbor df lives in the rows of the cells nc + k of the state buffers)
  int ib_ns = 0, ib_nf = 0;
  int64_t ib_rows = 0;               // velocity rows of all SolidNeighbors (after the cells')
  std::vector<int64_t> ib_donor;     // donor cell of each IB face
  bool ib_done = false;
  std::vector<int64_t> ib_of_pair;   // IB face of each pair, or -1
  DBuf d_ib_solid, d_ib_snoff, d_ib_snb, d_ib_faces, d_ib_img, d_ib_cut, d_ib_cutw, d_ib_ssdf;
  // velocity-space adaptation: the root lattice of the velocity grids (valid when vs_ok)
  VsArgs vsa{};
  bool vs_ok = false;
  bool ov_enable = true;   // overlap maps for adapted grids
  int64_t part_rows = 0;
  int cur = 0;
  bool timing = false;
  // waves per workgroup of the flux/relaxation kernel; 0: by shape. The FP32 form on adapted
  // grids whose LDS copy of f* leaves room for four or more workgroups per CU runs 4-wave
  // workgroups (more of them resident: measured 1.4x on 2D grids of 1.6k cells), and 2-wave
  // workgroups for the smallest size class (grids of up to 640 velocity cells: -7 % on the
  // flux phase of the adapted cylinder, whose free stream has 568-cell grids); larger grids,
  // whose LDS already limits residency, the 3D grids (8 waves: -14 % on the sphere's flux
  // phase against 4) and the other forms run 8 waves.
  int fr_waves = 0;
  int flux_waves(size_t lds) const {
    if (fr_waves) return fr_waves;
    if (D != 2 || !f32 || uni || lds == 0) return 8;   // 3D grids (4.5k cells and up): 8 waves
    if (lds <= 8 * 1024) return 2;
    return lds <= 40 * 1024 ? 4 : 8;
  }
  // launches of the flux/relaxation kernel: the fluid cells by grid size, each launch with
  // the LDS its largest grid needs (0: the global scratch area)
  struct FrBin {
    DBuf cells;
    int n = 0;
    size_t lds = 0;
    int waves = 8;          // waves per workgroup
    int reg_class = 0;      // fused step through the gather of regular cells (1, 2) or not
  };
  int reg_mode = 1;   // KAH_OPT_REGULAR: 0 every cell through the general slope and gather
  // KAH_OPT_BRANCHES: independent launches of one phase (the slope bins of one projection
  // depth, the flux/relaxation bins) spread over this many streams, forked from and joined to
  // the context's stream by events (parallel branches of the captured graph), so that small
  // launches run beside large ones instead of after them; 1: in order on the context's stream
  static constexpr int MAX_BRANCHES = 4;
  int branches = MAX_BRANCHES;   // measured: 7.95 -> 7.35 ms per step on the adapted cylinder
  hipStream_t side[MAX_BRANCHES - 1] = {};
  std::vector<hipEvent_t> ev_pool;   // fork/join events, reused step after step
  size_t ev_next = 0;
  std::vector<FrBin*> frbins;
  // velocity cells per thread of the chunked slope kernel; the default by dimension (set in
  // kah_create): 3 in 2D, 6 in 3D (measured on the adapted cylinder and sphere and the refined
  // Riemann mesh: 2D kernels above 4 need more registers, the 3D ones do not)
  int slope_vpt = 3;
  bool vt_enable = true;   // wave-tile order of shared tensor-product velocity grids
  int vt_mode = 0, vt_n0 = 0, vt_n1 = 0;
  std::vector<EvPair> events;
  // graph replay of the fused step: one executable graph per ping-pong parity, captured for
  // the step size dt it was recorded with
  bool use_graph = false;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  double gdt[2] = {0.0, 0.0};
  void drop_graphs() {
    for (int k = 0; k < 2; ++k) {
      if (gexec[k]) (void)hipGraphExecDestroy(gexec[k]);
      gexec[k] = nullptr;
    }
  }
  double t_sum[2] = {0, 0};
  int64_t t_cnt[2] = {0, 0};

  ~kah_ctx() {
    drop_graphs();
    for (auto* b : bins) delete b;
    for (auto* b : frbins) delete b;
    for (auto& e : events) {
      (void)hipEventDestroy(e.a);
      (void)hipEventDestroy(e.b);
    }
    for (hipEvent_t ev : ev_pool) (void)hipEventDestroy(ev);
    for (hipStream_t st : side)
      if (st) (void)hipStreamDestroy(st);
    if (stream) (void)hipStreamDestroy(stream);
  }
  // device memory held by the context
  int64_t device_bytes() const {
    const DBuf* bufs[] = {&d_fluid, &d_voff, &d_gu, &d_gw, &d_glev, &d_gkey, &d_gof, &d_df[0], &d_df[1],
                          &d_sdf, &d_oflux, &d_vflux, &d_psflux, &d_w, &d_prim, &d_qf, &d_half, &d_rho,
                          &d_res, &d_scratch, &d_cpre, &d_sl, &d_slnb, &d_hoff, &d_halves, &d_eoff,
                          &d_ents, &d_recds, &d_bc, &d_sw, &d_poff, &d_pcnt, &d_vtb, &d_ovmap,
                          &d_moff, &d_cmid, &d_benc, &d_xcells, &d_xoff, &d_xbuf, &d_xfluid,
                          &d_ib_solid, &d_ib_snoff, &d_ib_snb, &d_ib_faces, &d_ib_img,
                          &d_ib_cut, &d_ib_cutw, &d_ib_ssdf};
    int64_t t = 0;
    for (const DBuf* b : bufs) t += int64_t(b->bytes);
    for (const LevelBin* b : bins) t += int64_t(b->items.bytes);
    for (const FrBin* b : frbins) t += int64_t(b->cells.bytes);
    return t;
  }
  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }
  int hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return 0;
    err = std::string(what) + ": " + hipGetErrorString(e);
    return KAH_EDEVICE;
  }
  int alloc(DBuf& b, size_t bytes, const char* what) {
    b.release();
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
      b.p = nullptr;
      err = std::string("hipMalloc(") + what + ", " + std::to_string(bytes) +
            " B): " + hipGetErrorString(e);
      return e == hipErrorOutOfMemory ? KAH_ENOMEM : KAH_EDEVICE;
    }
    b.bytes = bytes;
    return 0;
  }
  // grow-only allocation (contents not kept)
  int reserve(DBuf& b, size_t bytes, const char* what) {
    if (b.p && b.bytes >= bytes) return 0;
    return alloc(b, bytes, what);
  }
  template <class T>
  int upload(DBuf& b, const T* src, size_t count, const char* what) {
    int rc = alloc(b, count * sizeof(T), what);
    if (rc) return rc;
    if (count == 0) return 0;
    return hip(hipMemcpy(b.p, src, count * sizeof(T), hipMemcpyHostToDevice), what);
  }
  size_t fsize() const { return f32 ? sizeof(float) : sizeof(double); }
  // host double array -> device storage type (float in the FP32 form: rounded to nearest)
  int upload_f(DBuf& b, const double* src, size_t count, const char* what) {
    if (!f32)
      return hip(hipMemcpyAsync(b.p, src, count * sizeof(double), hipMemcpyHostToDevice, stream),
                 what);
    std::vector<float> tmp(count);
    for (size_t x = 0; x < count; ++x) tmp[x] = float(src[x]);
    int rc = hip(hipMemcpyAsync(b.p, tmp.data(), count * sizeof(float), hipMemcpyHostToDevice,
                                stream), what);
    if (rc) return rc;
    return hip(hipStreamSynchronize(stream), what);
  }
  int download_f(double* dst, const DBuf& b, size_t count, const char* what) {
    return download_f(dst, b.p, count, what);
  }
  int download_f(double* dst, const void* src, size_t count, const char* what) {
    if (!dst) return 0;
    if (!f32)
      return hip(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToHost, stream),
                 what);
    std::vector<float> tmp(count);
    int rc = hip(hipMemcpyAsync(tmp.data(), src, count * sizeof(float), hipMemcpyDeviceToHost,
                                stream), what);
    if (!rc) rc = hip(hipStreamSynchronize(stream), what);
    if (rc) return rc;
    for (size_t x = 0; x < count; ++x) dst[x] = double(tmp[x]);
    return 0;
  }
  int zero(DBuf& b) {
    if (!b.p) return 0;
    return hip(hipMemsetAsync(b.p, 0, b.bytes, stream), "hipMemsetAsync");
  }
};

namespace {

#define KAH_TRY(x)            \
  do {                        \
    int _rc = (x);            \
    if (_rc) return _rc;      \
  } while (0)

// Key of a velocity cell (src/Velocity_space/Neighbor.jl:61-124): root lattice index times
// 2^(D*Lmax) plus the Morton code of the cell's lower corner inside its root.
uint64_t cell_key(int D, int L, int lmax, const double* mid, const double* lo,
                  const double* root_ds, const int32_t* ntree) {
  uint64_t root_lin = 0, stride = 1, morton = 0;
  for (int d = 0; d < D; ++d) {
    double hf = root_ds[d] / std::ldexp(1.0, lmax);
    double cell = hf * double(int64_t(1) << (lmax - L));
    int64_t g = std::llround((mid[d] - 0.5 * cell - lo[d]) / hf);
    root_lin += uint64_t(g >> lmax) * stride;
    stride *= uint64_t(ntree[d]);
    int64_t q = g & ((int64_t(1) << lmax) - 1);
    for (int b = 0; b < lmax; ++b) morton |= uint64_t((q >> b) & 1) << (b * D + d);
  }
  return (root_lin << (D * lmax)) | morton;
}

struct Builder {
  kah_ctx& C;
  const kah_topology& T;
  int D, NW;
  int64_t nc, nrec;
  std::vector<SlopeDesc> sl;
  std::vector<SlopeNb> slnb;
  std::vector<std::vector<HalfDesc>> hal;
  std::vector<std::vector<EntryDesc>> ent;

  Builder(kah_ctx& c, const kah_topology& t) : C(c), T(t), D(c.D), NW(c.NW) {
    nc = t.n_cells;
    nrec = t.n_cells + t.n_images;
  }
  int64_t src(int64_t rec) const { return rec < nc ? rec : T.image_src[rec - nc]; }
  const double* mid(int64_t rec) const { return &T.cell_mid[rec * D]; }
  const double* dsr(int64_t rec) const { return &T.cell_ds[rec * D]; }
  int32_t benc(int64_t rec) const { return T.bound_enc[src(rec)]; }
  int state(int64_t c, int f) const { return T.nb_state[c * 2 * D + f]; }
  int64_t nbn(int64_t c, int f) const {
    return T.nb_offset[c * 2 * D + f + 1] - T.nb_offset[c * 2 * D + f];
  }
  int64_t nb(int64_t c, int f, int64_t j) const { return T.nb_cell[T.nb_offset[c * 2 * D + f] + j]; }

  void add_side(SlopeDesc& s, int side, int64_t c, int f, int dir, bool proj) {
    s.nb_beg[side] = int32_t(slnb.size());
    s.nb_cnt[side] = int32_t(nbn(c, f));
    for (int64_t j = 0; j < nbn(c, f); ++j) {
      SlopeNb e{};
      e.cell = int32_t(nb(c, f, j));
      e.proj = proj ? 1 : 0;
      e.map = -1;
      for (int t = 0; t < 3; ++t) e.dm[t] = 0.0;
      if (proj)
        for (int t = 0; t < D; ++t) e.dm[t] = (t == dir) ? 0.0 : (mid(c)[t] - mid(e.cell)[t]);
      slnb.push_back(e);
    }
  }
  // _ps_has_transverse_offset, src/Flux/Slope.jl:177-199
  bool offset(int64_t c, int f, int dir) const {
    int64_t n = nbn(c, f);
    if (n == 0) return false;
    for (int t = 0; t < D; ++t) {
      if (t == dir) continue;
      double avg = 0.0;
      for (int64_t j = 0; j < n; ++j) avg += mid(nb(c, f, j))[t];
      avg /= double(n);
      if (avg != mid(c)[t]) return true;
    }
    return false;
  }
  static double inv_count(int32_t n) { return n > 0 ? 1.0 / double(n) : 0.0; }
  void inner(SlopeDesc& s, int64_t c, int dir, double dsL, double dsR, bool pL, bool pR) {
    s.mode = SL_INNER;
    s.ds[0] = dsL;
    s.ds[1] = dsR;
    s.rds[0] = 1.0 / dsL;
    s.rds[1] = 1.0 / dsR;
    add_side(s, 0, c, 2 * dir, dir, pL);
    add_side(s, 1, c, 2 * dir + 1, dir, pR);
    s.inv_cnt[0] = inv_count(s.nb_cnt[0]);
    s.inv_cnt[1] = inv_count(s.nb_cnt[1]);
  }
  void bound(SlopeDesc& s, int64_t c, int f, int dir, double ds, bool proj) {
    s.mode = SL_BOUND;
    s.ds[0] = ds;
    s.ds[1] = 0.0;
    s.rds[0] = 1.0 / ds;
    s.rds[1] = 0.0;
    add_side(s, 0, c, f, dir, proj);
    s.nb_beg[1] = 0;
    s.nb_cnt[1] = 0;
    s.inv_cnt[0] = inv_count(s.nb_cnt[0]);
    s.inv_cnt[1] = 0.0;
  }

  // update_slope!(Val(sL), Val(sR), ...), src/Flux/Slope.jl:458-771
  int plain(int64_t c, int dir, SlopeDesc& s) {
    int fL = 2 * dir, fR = 2 * dir + 1;
    int sL = state(c, fL), sR = state(c, fR);
    double ds = dsr(c)[dir], xm = mid(c)[dir];
    int F = 1 << (D - 1);
    bool FL = sL == F, FR = sR == F;
    if (sL == 1 && sR == 1) {
      bool Ls = benc(nb(c, fL, 0)) < 0, Rs = benc(nb(c, fR, 0)) < 0;
      if (Ls && Rs) s.mode = SL_ZERO;
      else if (Ls) bound(s, c, fR, dir, xm - mid(nb(c, fR, 0))[dir], false);
      else if (Rs) bound(s, c, fL, dir, xm - mid(nb(c, fL, 0))[dir], false);
      else inner(s, c, dir, ds, -ds, false, false);
    } else if (sL == 1 && FR) inner(s, c, dir, ds, -0.75 * ds, false, false);
    else if (FL && sR == 1) inner(s, c, dir, 0.75 * ds, -ds, false, false);
    else if (FL && FR) inner(s, c, dir, 0.75 * ds, -0.75 * ds, false, false);
    else if (sL == -1 && sR == -1) inner(s, c, dir, 1.5 * ds, -1.5 * ds, false, false);
    else if (sL == 1 && sR == -1) inner(s, c, dir, ds, -1.5 * ds, false, false);
    else if (sL == -1 && sR == 1) inner(s, c, dir, 1.5 * ds, -ds, false, false);
    else if (FL && sR == -1) inner(s, c, dir, 0.75 * ds, -1.5 * ds, false, false);
    else if (sL == -1 && FR) inner(s, c, dir, 1.5 * ds, -0.75 * ds, false, false);
    else if (sL == 0 && sR != 0) bound(s, c, fR, dir, xm - mid(nb(c, fR, 0))[dir], false);
    else if (sR == 0 && sL != 0) bound(s, c, fL, dir, xm - mid(nb(c, fL, 0))[dir], false);
    else
      return C.fail(KAH_EINVAL, "update_slope!: no method for neighbour states (" +
                                    std::to_string(sL) + "," + std::to_string(sR) +
                                    ") of cell " + std::to_string(c));
    return 0;
  }
  // per-direction outcome of update_slope_transverse_level!, src/Flux/Slope.jl:849-945
  int transverse(int64_t c, int dir, SlopeDesc& s) {
    int fL = 2 * dir, fR = 2 * dir + 1;
    int sL = state(c, fL), sR = state(c, fR);
    double xm = mid(c)[dir];
    if (sL != 0 && sR != 0) {
      int64_t L0 = nb(c, fL, 0), R0 = nb(c, fR, 0);
      double dsL = xm - mid(L0)[dir], dsR = xm - mid(R0)[dir];
      if (benc(L0) < 0) bound(s, c, fR, dir, dsR, true);
      else if (benc(R0) < 0) bound(s, c, fL, dir, dsL, true);
      else inner(s, c, dir, dsL, dsR, offset(c, fL, dir), offset(c, fR, dir));
    } else if (sR == 0 && sL == -1) {
      int64_t L0 = nb(c, fL, 0);
      if (benc(L0) < 0) s.mode = SL_SKIP;
      else bound(s, c, fL, dir, xm - mid(L0)[dir], true);
    } else {  // sL == 0 && sR == -1
      int64_t R0 = nb(c, fR, 0);
      if (benc(R0) < 0) s.mode = SL_SKIP;
      else bound(s, c, fR, dir, xm - mid(R0)[dir], true);
    }
    return 0;
  }

  HalfDesc half(int32_t slot, int32_t data, int8_t mode, int8_t sel, int dir, double rot,
                int64_t rec, const double* fm, int32_t bc) {
    HalfDesc h{};
    h.slot = slot;
    h.data_cell = data;
    h.mode = mode;
    h.sel = sel;
    h.dir = int8_t(dir);
    h.bc = bc;
    h.rot = rot;
    for (int t = 0; t < D; ++t) {
      h.xrec[t] = mid(rec)[t];
      h.dsrec[t] = dsr(rec)[t];
      h.fmid[t] = fm[t];
      h.dx_f[t] = float(h.fmid[t] - h.xrec[t]);
      h.ds_f[t] = float(h.dsrec[t]);
    }
    h.rot_f = float(rot);
    return h;
  }
  EntryDesc entry(int32_t pair, int8_t side, int8_t om, int8_t tm, int32_t oth, int dir,
                  double rot, double area, const double* fm, int64_t own_rec, int64_t oth_rec,
                  int32_t bc) {
    EntryDesc e{};
    e.pair = pair;
    e.side = side;
    e.own_mode = om;
    e.oth_mode = tm;
    e.oth_cell = oth;
    e.dir = int8_t(dir);
    e.bc = bc;
    e.rot = rot;
    e.area = area;
    for (int t = 0; t < D; ++t) {
      e.fmid[t] = fm[t];
      e.xown[t] = mid(own_rec)[t];
      e.dsown[t] = dsr(own_rec)[t];
      e.xoth[t] = oth_rec >= 0 ? mid(oth_rec)[t] : 0.0;
      e.dsoth[t] = oth_rec >= 0 ? dsr(oth_rec)[t] : 0.0;
      e.dxo_f[t] = float(e.fmid[t] - e.xown[t]);
      e.dso_f[t] = float(e.dsown[t]);
      e.dxt_f[t] = float(e.fmid[t] - e.xoth[t]);
      e.dst_f[t] = float(e.dsoth[t]);
    }
    e.rot_f = float(rot);
    e.area_f = float(area);
    return e;
  }

  int build() {
    const bool caidvm = C.P.flux_kind == KAH_FLUX_CAIDVM;
    sl.assign(size_t(nc * D), SlopeDesc{});
    hal.assign(size_t(nc), {});
    ent.assign(size_t(nc), {});
    int Lmin = 1 << 30;
    for (int64_t c = 0; c < nc; ++c)
      if (T.bound_enc[c] >= 0) Lmin = std::min(Lmin, int(T.ps_level[c]));
    for (int64_t c = 0; c < nc; ++c) {
      if (T.bound_enc[c] < 0) continue;
      bool fine = T.ps_level[c] > Lmin;
      for (int d = 0; d < D; ++d) {
        SlopeDesc& s = sl[c * D + d];
        s.mode = SL_ZERO;
        int sL = state(c, 2 * d), sR = state(c, 2 * d + 1);
        if (fine && D > 1 && (sL == -1 || sR == -1)) KAH_TRY(transverse(c, d, s));
        else KAH_TRY(plain(c, d, s));
      }
    }
    for (int64_t p = 0; p < T.n_pairs; ++p) {
      int64_t h = T.pair_here[p], trec = T.pair_there[p];
      if (h < 0 || h >= nc || T.bound_enc[h] < 0)
        return C.fail(KAH_EINVAL, "pair " + std::to_string(p) + ": here is not a local fluid cell");
      int dir = T.pair_dir[p];
      double rot = T.pair_rot[p], area = T.pair_area[p];
      const double* fm = &T.pair_mid[p * D];
      int32_t s0 = int32_t(2 * p), s1 = int32_t(2 * p + 1);
      if (trec < 0) {
        int32_t bc = T.pair_bc[p];
        if (bc < 0 || bc >= T.n_bc) return C.fail(KAH_EINVAL, "domain pair without a boundary condition");
        int kind = T.bc_kind[bc];
        int8_t m0, m1, mo;
        switch (kind) {
          case KAH_BC_UNIFORM_OUTFLOW: m0 = MI_PLAIN; m1 = MI_PLAIN; mo = MI_SELF_PLAIN; break;
          case KAH_BC_SUPERSONIC_INFLOW: m0 = MI_UNLIMITED; m1 = MI_MAXWELL; mo = MI_MAXWELL; break;
          case KAH_BC_MAXWELLIAN:
            if (!caidvm)
              return C.fail(KAH_EUNSUPPORTED,
                            "UndefVarError: `here_weight` not defined (DVM Maxwellian domain "
                            "flux, src/Flux/DVM.jl:12)");
            m0 = MI_UNLIMITED; m1 = MI_WALL; mo = MI_WALL; break;
          case KAH_BC_INTERPOLATED_OUTFLOW:
            if (D != 2)
              return C.fail(KAH_EUNSUPPORTED, "InterpolatedOutflow is defined for DIM=2 only");
            m0 = MI_UNLIMITED; m1 = MI_INTERP; mo = MI_INTERP; break;
          default:
            return C.fail(KAH_EINVAL, "unknown boundary condition kind " + std::to_string(kind));
        }
        if (caidvm) {
          hal[h].push_back(half(s0, int32_t(h), m0, 0, dir, rot, h, fm, bc));
          hal[h].push_back(half(s1, int32_t(h), m1, 1, dir, rot, h, fm, bc));
        }
        ent[h].push_back(entry(int32_t(p), 0, m0, mo, -1, dir, rot, area, fm, h, -1, bc));
      } else {
        if (trec >= nrec) return C.fail(KAH_EINVAL, "pair there out of range");
        int64_t t = src(trec);
        bool solid = T.bound_enc[t] < 0;
        // an IB face reads its there side from the SolidNeighbor rows (cell nc + k), whose
        // half flux the IB kernel writes after the slopes
        const int64_t ibk = C.ib_of_pair.empty() ? -1 : C.ib_of_pair[size_t(p)];
        if (solid && !C.uni && ibk < 0)
          return C.fail(KAH_EUNSUPPORTED,
                        "faces with solid cells outside an immersed boundary need a shared "
                        "velocity grid");
        int8_t mh = solid ? MI_UNLIMITED : MI_LIMITED;
        int8_t mt = solid ? (caidvm ? MI_PLAIN : MI_UNLIMITED) : MI_LIMITED;
        if (caidvm) {
          hal[h].push_back(half(s0, int32_t(h), mh, 0, dir, rot, h, fm, -1));
          if (solid && ibk < 0)
            hal[h].push_back(half(s1, int32_t(t), MI_PLAIN, 1, dir, rot, trec, fm, -1));
          else if (!solid)
            hal[t].push_back(half(s1, int32_t(t), MI_LIMITED, 1, dir, rot, trec, fm, -1));
        }
        const int32_t oth = ibk >= 0 ? int32_t(nc + ibk) : int32_t(t);
        ent[h].push_back(entry(int32_t(p), 0, mh, mt, oth, dir, rot, area, fm, h, trec, -1));
        if (!solid)
          ent[t].push_back(entry(int32_t(p), 1, MI_LIMITED, MI_LIMITED, int32_t(h), dir, rot, area,
                                 fm, trec, h, -1));
      }
    }
    return 0;
  }
};

// Immersed-boundary arrays of the topology: checked, then uploaded as device descriptors.
// Fills C->ib_of_pair for the descriptor builder.
int setup_ib(kah_ctx* C, const kah_topology* T) {
  const int D = C->D, NW = C->NW;
  const int64_t nc = T->n_cells;
  C->ib_ns = 0;
  C->ib_nf = 0;
  C->ib_rows = 0;
  C->ib_donor.clear();
  C->ib_done = false;
  C->ib_of_pair.clear();
  for (DBuf* b : {&C->d_ib_solid, &C->d_ib_snoff, &C->d_ib_snb, &C->d_ib_faces, &C->d_ib_img,
                  &C->d_ib_cut, &C->d_ib_cutw, &C->d_ib_ssdf})
    b->release();
  const int64_t ns = T->n_ib_solid, nf = T->n_ib_faces;
  if (ns == 0 && nf == 0) return 0;
  if (ns < 0 || nf < 0 || ns > nc || nf > T->n_pairs)
    return C->fail(KAH_EINVAL, "immersed boundary counts out of range");
  if ((ns > 0 && (!T->ib_solid || !T->ib_solid_nb_off || !T->ib_solid_nb)) ||
      (nf > 0 && (!T->ib_face_pair || !T->ib_aux_point || !T->ib_normal || !T->ib_bc ||
                  !T->ib_image_off || !T->ib_image_cell || !T->ib_cut_off ||
                  (T->ib_cut_off[nf] > 0 &&
                   (!T->ib_cut_index || !T->ib_cut_gas || !T->ib_cut_solid)))))
    return C->fail(KAH_EINVAL, "null immersed-boundary array");
  auto fluid = [&](int64_t c) { return c >= 0 && c < nc && T->bound_enc[c] >= 0; };
  std::vector<int32_t> sol(static_cast<size_t>(ns)), soff(static_cast<size_t>(ns + 1), 0), snb;
  for (int64_t q = 0; q < ns; ++q) {
    const int64_t s = T->ib_solid[q];
    if (s < 0 || s >= nc || T->bound_enc[s] >= 0)
      return C->fail(KAH_EINVAL, "ib_solid[" + std::to_string(q) + "] is not a solid cell");
    const int64_t b = T->ib_solid_nb_off[q], e = T->ib_solid_nb_off[q + 1];
    if (b != int64_t(snb.size()) || e - b < 1 || e - b > 8)
      return C->fail(KAH_EINVAL, "solid cell " + std::to_string(s) +
                                     ": 1 to 8 fluid neighbours in order expected");
    for (int64_t x = b; x < e; ++x) {
      if (!fluid(T->ib_solid_nb[x]))
        return C->fail(KAH_EINVAL, "solid cell " + std::to_string(s) + ": neighbour " +
                                       std::to_string(T->ib_solid_nb[x]) + " is not fluid");
      snb.push_back(int32_t(T->ib_solid_nb[x]));
    }
    sol[q] = int32_t(s);
    soff[q + 1] = int32_t(snb.size());
  }
  C->ib_of_pair.assign(size_t(T->n_pairs), -1);
  std::vector<IbFaceDesc> fd(static_cast<size_t>(nf));
  std::vector<int32_t> img, cut;
  std::vector<double2> cutw;
  for (int64_t k = 0; k < nf; ++k) {
    const int64_t p = T->ib_face_pair[k];
    if (p < 0 || p >= T->n_pairs || C->ib_of_pair[size_t(p)] >= 0)
      return C->fail(KAH_EINVAL, "ib_face_pair[" + std::to_string(k) + "] out of range or repeated");
    const int64_t h = T->pair_here[p], t = T->pair_there[p];
    if (h < 0 || h >= nc || T->bound_enc[h] <= 0 || t < 0 || t >= nc || T->bound_enc[t] >= 0)
      return C->fail(KAH_EINVAL, "IB face " + std::to_string(k) +
                                     ": pair must join a donor (here) and a solid cell (there)");
    C->ib_of_pair[size_t(p)] = k;
    C->ib_donor.push_back(h);
    const int64_t nd = T->vs_offset[h + 1] - T->vs_offset[h];
    IbFaceDesc& F = fd[size_t(k)];
    F.row = C->ib_rows;
    C->ib_rows += nd;
    F.donor = int32_t(h);
    F.solid = int32_t(t);
    F.pseudo = int32_t(nc + k);
    F.dir = T->pair_dir[p];
    F.slot = int32_t(2 * p + 1);
    F.rot = T->pair_rot[p];
    const int64_t ib = T->ib_image_off[k], ie = T->ib_image_off[k + 1];
    if (ib != int64_t(img.size()) || ie - ib < 1 || ie - ib > 8)
      return C->fail(KAH_EINVAL, "IB face " + std::to_string(k) + ": 1 to 8 image cells in order expected");
    for (int64_t x = ib; x < ie; ++x) {
      if (!fluid(T->ib_image_cell[x]))
        return C->fail(KAH_EINVAL, "IB face " + std::to_string(k) + ": image cell " +
                                       std::to_string(T->ib_image_cell[x]) + " is not fluid");
      img.push_back(int32_t(T->ib_image_cell[x]));
    }
    F.img_beg = int32_t(ib);
    F.img_cnt = int32_t(ie - ib);
    const int64_t cb = T->ib_cut_off[k], ce = T->ib_cut_off[k + 1];
    if (cb != int64_t(cut.size()) || ce < cb)
      return C->fail(KAH_EINVAL, "IB face " + std::to_string(k) + ": cut offsets out of order");
    for (int64_t x = cb; x < ce; ++x) {
      const int32_t i = T->ib_cut_index[x];
      if (i < 0 || i >= nd || (x > cb && i <= T->ib_cut_index[x - 1]) ||
          !(T->ib_cut_gas[x] > 0.0) || !(T->ib_cut_solid[x] > 0.0))
        return C->fail(KAH_EINVAL, "IB face " + std::to_string(k) +
                                       ": cut cells must be ascending, in range, with positive areas");
      cut.push_back(i);
      cutw.push_back(make_double2(T->ib_cut_gas[x], T->ib_cut_solid[x]));
    }
    F.cut_beg = int32_t(cb);
    F.cut_cnt = int32_t(ce - cb);
    const double* ap = &T->ib_aux_point[k * D];
    const double* xd = &T->cell_mid[h * D];
    const double* xs = &T->cell_mid[t * D];
    for (int d = 0; d < 3; ++d) {
      F.ip[d] = d < D ? ap[d] + xd[d] - xs[d] : 0.0;
      F.nrm[d] = d < D ? T->ib_normal[k * D + d] : 0.0;
    }
    F.dxs = xd[F.dir] - xs[F.dir];
    F.dxf = F.ip[F.dir] - xd[F.dir];
    F.dxL = ap[F.dir] - F.ip[F.dir];
    F.apd = ap[F.dir];
    for (int m = 0; m < 5; ++m) F.bc[m] = m < NW ? T->ib_bc[k * NW + m] : 0.0;
    F.bc[0] = 1.0;
  }
  C->ib_ns = int(ns);
  C->ib_nf = int(nf);
  if (ns > 0) {
    KAH_TRY(C->upload(C->d_ib_solid, sol.data(), sol.size(), "ib solid cells"));
    KAH_TRY(C->upload(C->d_ib_snoff, soff.data(), soff.size(), "ib solid neighbour offsets"));
    KAH_TRY(C->upload(C->d_ib_snb, snb.data(), snb.size(), "ib solid neighbours"));
  }
  if (nf > 0 && C->P.marching == KAH_MARCH_CIP)   // the limited boundary slopes
    KAH_TRY(C->alloc(C->d_ib_ssdf, sizeof(double) * size_t(C->ib_rows) * size_t(C->NDF),
                     "ib boundary slopes"));
  if (nf > 0) {
    KAH_TRY(C->upload(C->d_ib_faces, fd.data(), fd.size(), "ib faces"));
    KAH_TRY(C->upload(C->d_ib_img, img.data(), img.size(), "ib image cells"));
    if (!cut.empty()) {
      KAH_TRY(C->upload(C->d_ib_cut, cut.data(), cut.size(), "ib cut cells"));
      KAH_TRY(C->upload(C->d_ib_cutw, cutw.data(), cutw.size(), "ib cut areas"));
    }
  }
  return 0;
}

}  // namespace

// ------------------------------------------------ scratch of the allocation-free entry point
// kah_vs_resolution_noalloc keeps its per-cell partial maxima (two per cell) and its two
// results in this fixed device array instead of memory it allocates: 16 MiB plus 16 bytes,
// held by every GPU the library runs on until the process ends. In host code `scratch` only
// names the symbol; its device address, one per GPU, comes from hipGetSymbolAddress in
// kah_create (with that GPU current) and is the only form host code passes on.
constexpr int64_t NOALLOC_MAX_CELLS = int64_t(1) << 20;
static __device__ __align__(256) double scratch[2 * NOALLOC_MAX_CELLS + 2];
namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
hipError_t g_scratch_err[NOALLOC_MAX_DEVICES] = {};
// once per device, with that device current
void scratch_setup(int dev) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess && p) {
    g_scratch_ptr[dev] = static_cast<double*>(p);
  } else {
    g_scratch_err[dev] = e == hipSuccess ? hipErrorInvalidSymbol : e;
    (void)hipGetLastError();
  }
}
}  // namespace

// =========================================================================== C ABI
extern "C" {

int kah_create(const kah_params* params, kah_ctx** out) {
  if (!out) return KAH_EINVAL;
  *out = nullptr;
  if (!params) return KAH_EINVAL;
  if (!((params->dim == 2 && params->ndf == 2) || (params->dim == 3 && params->ndf == 1)))
    return KAH_EUNSUPPORTED;
  if (params->flux_kind != KAH_FLUX_CAIDVM && params->flux_kind != KAH_FLUX_DVM)
    return KAH_EINVAL;
  if (params->marching != KAH_MARCH_CAIDVM && params->marching != KAH_MARCH_EULER &&
      params->marching != KAH_MARCH_CIP)
    return KAH_EINVAL;
  if (params->precision != KAH_PREC_FP64 && params->precision != KAH_PREC_FP32F)
    return KAH_EUNSUPPORTED;
  kah_ctx* C = new kah_ctx();
  C->P = *params;
  if (C->P.st_check_interval <= 0) C->P.st_check_interval = 100;
  C->D = params->dim;
  C->NDF = params->ndf;
  C->NW = params->dim + 2;
  C->slope_vpt = params->dim == 3 ? 6 : 3;
  C->f32 = params->precision == KAH_PREC_FP32F;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    delete C;
    return KAH_EDEVICE;
  }
  if (params->device < 0 || params->device >= ndev || hipSetDevice(params->device) != hipSuccess ||
      hipStreamCreate(&C->stream) != hipSuccess) {
    delete C;
    return KAH_EDEVICE;
  }
  for (hipStream_t& st : C->side)   // default priority, like the context's stream
    if (hipStreamCreate(&st) != hipSuccess) {
      delete C;
      return KAH_EDEVICE;
    }
  scratch_setup(params->device);   // the allocation-free entry point's scratch on this GPU
  // dynamic LDS per workgroup: the device's limit (160 KiB on gfx950) less the kernels'
  // static LDS
  int lds_max = 0;
  if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock,
                            params->device) == hipSuccess && lds_max > 16 * 1024)
    C->lds_limit = size_t(lds_max) - 8 * 1024;
  *out = C;
  return KAH_OK;
}

void kah_destroy(kah_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->P.device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
}

const char* kah_last_error(const kah_ctx* ctx) {
  if (!ctx) return "null kah_ctx";
  return ctx->err.c_str();
}

int kah_set_topology(kah_ctx* C, const kah_topology* T) {
  if (!C) return KAH_EINVAL;
  C->drop_graphs();   // captured launches refer to the buffers replaced below
  if (!T) return C->fail(KAH_EINVAL, "null topology");
  KAH_TRY(C->hip(hipSetDevice(C->P.device), "hipSetDevice"));
  C->have_topo = C->have_state = false;
  const int D = C->D;
  const int64_t nc = T->n_cells;
  if (nc <= 0) return C->fail(KAH_EINVAL, "n_cells must be positive");
  if (nc >= (int64_t(1) << 31)) return C->fail(KAH_EUNSUPPORTED, "more than 2^31 cells");
  if (!T->cell_mid || !T->cell_ds || !T->bound_enc || !T->ps_level || !T->vs_offset ||
      !T->vs_level || !T->vs_mid || !T->vs_weight || !T->nb_state || !T->nb_offset ||
      (T->n_images > 0 && !T->image_src) ||
      (T->n_pairs > 0 && (!T->pair_here || !T->pair_there || !T->pair_dir || !T->pair_rot ||
                          !T->pair_area || !T->pair_mid || !T->pair_bc)))
    return C->fail(KAH_EINVAL, "null array in topology");
  for (int64_t c = 0; c < nc; ++c)
    if (T->vs_offset[c + 1] <= T->vs_offset[c])
      return C->fail(KAH_EINVAL, "empty or unordered velocity grid of cell " + std::to_string(c));
  C->nc = nc;
  C->np = T->vs_offset[nc];
  C->npairs = T->n_pairs;
  C->uni = T->vs_uniform != 0;
  C->n0 = int(T->vs_offset[1] - T->vs_offset[0]);
  C->lmax = C->uni ? 0 : T->vs_maxlevel;
  C->max_n = 0;
  for (int64_t c = 0; c < nc; ++c) {
    int64_t n = T->vs_offset[c + 1] - T->vs_offset[c];
    if (n >= (int64_t(1) << 31)) return C->fail(KAH_EUNSUPPORTED, "velocity grid too large");
    C->max_n = std::max<int>(C->max_n, int(n));
    if (C->uni && T->vs_offset[c] != c * int64_t(C->n0))
      return C->fail(KAH_EINVAL, "vs_uniform requires vs_offset[c] == c*n0");
  }
  for (int64_t k = 0; k < T->n_images; ++k)
    if (T->image_src[k] < 0 || T->image_src[k] >= nc)
      return C->fail(KAH_EINVAL, "image source out of range");
  const int64_t nb_total = T->nb_offset[nc * 2 * D];
  for (int64_t x = 0; x < nb_total; ++x)
    if (T->nb_cell[x] < 0 || T->nb_cell[x] >= nc)
      return C->fail(KAH_EINVAL, "neighbour id out of range");
  C->benc.assign(T->bound_enc, T->bound_enc + nc);

  // Wave tiles of a shared tensor-product velocity grid (vs_trees_num cells per direction,
  // first direction fastest): 8x8 (2D) or 4x4x4 (3D) blocks, so that the 64 lanes of a wave
  // share the sign of every velocity component when the zero planes fall on tile
  // boundaries, and upwind branches do not split waves. Any tile order is a bijection of
  // the grid, so results do not depend on it.
  C->vt_mode = 0;
  C->d_vtb.release();
  if (C->uni && C->vt_enable) {
    const int tw = D == 2 ? 8 : 4;
    int64_t prod = 1;
    bool ok = C->n0 % 64 == 0;
    for (int d = 0; d < D; ++d) {
      ok = ok && T->vs_trees_num[d] > 0 && T->vs_trees_num[d] % tw == 0;
      prod *= std::max(T->vs_trees_num[d], 1);
    }
    if (ok && prod == C->n0) {
      const int n0 = T->vs_trees_num[0], n1 = T->vs_trees_num[1];
      const int tx = n0 / tw, ty = n1 / tw, tz = D == 3 ? T->vs_trees_num[2] / tw : 1;
      std::vector<int32_t> base;
      base.reserve(size_t(tx) * ty * tz);
      for (int z = 0; z < tz; ++z)
        for (int y = 0; y < ty; ++y)
          for (int x = 0; x < tx; ++x)
            base.push_back(int32_t((int64_t(z) * tw * n1 + int64_t(y) * tw) * n0 + x * tw));
      KAH_TRY(C->upload(C->d_vtb, base.data(), base.size(), "velocity tiles"));
      C->vt_mode = D;
      C->vt_n0 = n0;
      C->vt_n1 = n1;
    }
  }

  // root lattice for velocity-space adaptation (VsNeighborIndex, src/Velocity_space/
  // Neighbor.jl:61-124: ds0 = quadrature extent / vs_trees_num, h_fine = ds0 / 2^maxlevel)
  C->vs_ok = T->vs_maxlevel >= 0 && T->vs_maxlevel < 16;
  C->vsa = VsArgs{};
  {
    double du = 1.0;
    int64_t tn = 1, stride = 1;
    for (int d = 0; d < D; ++d) {
      if (T->vs_trees_num[d] <= 0 || !(T->vs_hi[d] > T->vs_lo[d])) C->vs_ok = false;
      if (!C->vs_ok) break;
      C->vsa.vmin[d] = T->vs_lo[d];
      C->vsa.hf[d] = ((T->vs_hi[d] - T->vs_lo[d]) / double(T->vs_trees_num[d])) /
                     std::ldexp(1.0, T->vs_maxlevel);
      C->vsa.tn[d] = T->vs_trees_num[d];
      C->vsa.stride[d] = stride;
      stride *= T->vs_trees_num[d];
      du = d == 0 ? (T->vs_hi[d] - T->vs_lo[d]) : du * (T->vs_hi[d] - T->vs_lo[d]);
      tn *= T->vs_trees_num[d];
    }
    C->vsa.L = T->vs_maxlevel;
    if (C->vs_ok) C->vsa.wfine = du / double(tn) / double(int64_t(1) << (D * T->vs_maxlevel));
  }

  // velocity grids: upload, and keys for adapted grids
  const int64_t ng = C->uni ? C->n0 : C->np;
  std::vector<uint32_t> keys;
  if (!C->uni) {
    int lmax = T->vs_maxlevel;
    double root_ds[3];
    int64_t roots = 1;
    for (int d = 0; d < D; ++d) {
      if (T->vs_trees_num[d] <= 0 || !(T->vs_hi[d] > T->vs_lo[d]))
        return C->fail(KAH_EINVAL, "velocity root grid (vs_trees_num / vs_lo / vs_hi) not set");
      root_ds[d] = 2.0 * ((T->vs_hi[d] - T->vs_lo[d]) / (2.0 * T->vs_trees_num[d]));
      roots *= T->vs_trees_num[d];
    }
    int rootbits = 0;
    while ((int64_t(1) << rootbits) < roots) ++rootbits;
    if (rootbits + D * lmax > 32)
      return C->fail(KAH_EUNSUPPORTED, "velocity cell keys exceed 32 bits");
    keys.resize(size_t(ng));
    for (int64_t c = 0; c < nc; ++c) {
      int64_t off = T->vs_offset[c], n = T->vs_offset[c + 1] - off;
      for (int64_t i = 0; i < n; ++i) {
        double m[3];
        for (int d = 0; d < D; ++d) m[d] = T->vs_mid[off * D + d * n + i];
        int L = T->vs_level[off + i];
        if (L < 0 || L > lmax) return C->fail(KAH_EINVAL, "velocity level out of range");
        uint64_t k = cell_key(D, L, lmax, m, T->vs_lo, root_ds, T->vs_trees_num);
        keys[size_t(off + i)] = uint32_t(k);
        if (i > 0 && keys[size_t(off + i)] <= keys[size_t(off + i - 1)])
          return C->fail(KAH_EINVAL, "velocity grid of cell " + std::to_string(c) +
                                         " is not in key order");
      }
    }
  }
  // immersed boundary: on adapted grids each SolidNeighbor (cell nc + k) reads its donor's
  // grid, so kernels read it like any cell
  KAH_TRY(setup_ib(C, T));
  // Distinct velocity grids (adapted grids): cells whose grids are bit-identical (levels,
  // midpoints, weights) share one copy of the geometry, and gof[c] is the first entry of cell
  // c's grid in that table. The state (f, sdf) stays per cell at vs_offset. A face neighbour
  // on the same grid then needs no overlap map (velocity cell i meets velocity cell i), and
  // cells of one grid read its geometry from the caches.
  std::vector<int64_t> gof;
  std::vector<int64_t> voff(T->vs_offset, T->vs_offset + nc + 1);
  for (int64_t h : C->ib_donor) voff.push_back(voff.back() + (T->vs_offset[h + 1] - T->vs_offset[h]));
  C->n_grids = 1;
  if (!C->uni) {
    std::vector<double> gu, gw;
    std::vector<int8_t> gl;
    std::vector<uint32_t> gk;
    gof.assign(size_t(nc) + C->ib_donor.size(), 0);
    // hash of a grid's bytes -> first cell of each distinct grid with that hash
    std::unordered_map<uint64_t, std::vector<int64_t>> seen;
    seen.reserve(size_t(nc));
    auto fnv = [](uint64_t h, const void* p, size_t n) {   // FNV-1a over 8-byte words
      const unsigned char* b = static_cast<const unsigned char*>(p);
      size_t x = 0;
      for (; x + 8 <= n; x += 8) {
        uint64_t w;
        std::memcpy(&w, b + x, 8);
        h = (h ^ w) * 1099511628211ull;
      }
      for (; x < n; ++x) h = (h ^ b[x]) * 1099511628211ull;
      return h;
    };
    auto same_grid = [&](int64_t a, int64_t b) {
      const int64_t oa = T->vs_offset[a], ob = T->vs_offset[b];
      const int64_t n = T->vs_offset[a + 1] - oa;
      if (T->vs_offset[b + 1] - ob != n) return false;
      return std::memcmp(T->vs_level + oa, T->vs_level + ob, size_t(n)) == 0 &&
             std::memcmp(T->vs_mid + oa * D, T->vs_mid + ob * D, sizeof(double) * n * D) == 0 &&
             std::memcmp(T->vs_weight + oa, T->vs_weight + ob, sizeof(double) * n) == 0;
    };
    std::vector<int64_t> rep_of(size_t(nc), -1);   // representative cell of each cell's grid
    int64_t ndistinct = 0;
    for (int64_t c = 0; c < nc; ++c) {
      const int64_t o = T->vs_offset[c], n = T->vs_offset[c + 1] - o;
      uint64_t h = 1469598103934665603ull;
      h = fnv(h, &n, sizeof(n));
      h = fnv(h, T->vs_level + o, size_t(n));
      h = fnv(h, keys.data() + o, sizeof(uint32_t) * size_t(n));
      auto& lst = seen[h];
      int64_t rep = -1;
      for (int64_t r : lst)
        if (same_grid(r, c)) {
          rep = r;
          break;
        }
      if (rep < 0) {
        lst.push_back(c);
        rep = c;
        gof[size_t(c)] = int64_t(gw.size());
        gu.insert(gu.end(), T->vs_mid + o * D, T->vs_mid + (o + n) * D);
        gw.insert(gw.end(), T->vs_weight + o, T->vs_weight + o + n);
        gl.insert(gl.end(), T->vs_level + o, T->vs_level + o + n);
        gk.insert(gk.end(), keys.begin() + o, keys.begin() + o + n);
        ++ndistinct;
      } else {
        gof[size_t(c)] = gof[size_t(rep)];
      }
      rep_of[size_t(c)] = rep;
    }
    for (size_t k = 0; k < C->ib_donor.size(); ++k)
      gof[size_t(nc) + k] = gof[size_t(C->ib_donor[k])];
    C->n_grids = ndistinct;
    C->grid_cells = int64_t(gw.size());
    KAH_TRY(C->upload(C->d_gu, gu.data(), gu.size(), "vs_mid"));
    KAH_TRY(C->upload(C->d_gw, gw.data(), gw.size(), "vs_weight"));
    KAH_TRY(C->upload(C->d_glev, gl.data(), gl.size(), "vs_level"));
    KAH_TRY(C->upload(C->d_gkey, gk.data(), gk.size(), "vs_key"));
    KAH_TRY(C->upload(C->d_gof, gof.data(), gof.size(), "grid offsets"));
  } else {
    C->grid_cells = C->n0;
    KAH_TRY(C->upload(C->d_gu, T->vs_mid, size_t(ng * D), "vs_mid"));
    KAH_TRY(C->upload(C->d_gw, T->vs_weight, size_t(ng), "vs_weight"));
    KAH_TRY(C->upload(C->d_glev, T->vs_level, size_t(ng), "vs_level"));
    C->d_gkey.release();
    C->d_gof.release();
  }
  KAH_TRY(C->upload(C->d_voff, voff.data(), voff.size(), "vs_offset"));

  // descriptors
  Builder B(*C, *T);
  KAH_TRY(B.build());

  // Overlap maps of adapted grids: one map per distinct ordered pair (owner grid, neighbour
  // grid) that a fluid cell's slope stencil or face gather reads across grids (see OvJob); a
  // neighbour on the owner's own grid reads velocity cell i itself (MAP_SAME).
  std::vector<int64_t> mbase;
  std::vector<OvJob> jobs;
  const bool maps = !C->uni && C->ov_enable && D * C->lmax <= 7;
  {
    std::unordered_map<uint64_t, int32_t> map_id;   // (owner grid, neighbour grid) -> map
    int64_t next = 0;
    auto assign = [&](int64_t c, int32_t nbc) -> int32_t {
      if (C->uni) return -1;
      const int64_t go = gof[size_t(c)], gn = gof[size_t(nbc)];
      if (go == gn) return MAP_SAME;
      if (!maps) return -1;
      const uint64_t key = uint64_t(go) * uint64_t(C->grid_cells + 1) + uint64_t(gn);
      auto it = map_id.find(key);
      if (it != map_id.end()) return it->second;
      const int32_t id = int32_t(mbase.size());
      const int32_t no = int32_t(voff[size_t(c) + 1] - voff[size_t(c)]);
      const int32_t nn = int32_t(voff[size_t(nbc) + 1] - voff[size_t(nbc)]);
      mbase.push_back(next);
      jobs.push_back(OvJob{go, gn, no, nn, next});
      next += no;
      map_id.emplace(key, id);
      return id;
    };
    for (auto& e : B.slnb) e.map = -1;
    for (int64_t c = 0; c < nc; ++c) {
      for (auto& ed : B.ent[c]) ed.map = -1;
      if (T->bound_enc[c] < 0) continue;
      for (int d = 0; d < D; ++d) {
        const SlopeDesc& sd = B.sl[size_t(c * D + d)];
        const int sides = sd.mode == SL_INNER ? 2 : (sd.mode == SL_BOUND ? 1 : 0);
        for (int sdx = 0; sdx < sides; ++sdx)
          for (int e = 0; e < sd.nb_cnt[sdx]; ++e) {
            SlopeNb& nbr = B.slnb[size_t(sd.nb_beg[sdx] + e)];
            nbr.map = assign(c, nbr.cell);
          }
      }
      for (auto& ed : B.ent[c])
        if (ed.oth_cell >= 0 &&
            (ed.oth_mode == MI_LIMITED || ed.oth_mode == MI_UNLIMITED || ed.oth_mode == MI_PLAIN))
          ed.map = assign(c, ed.oth_cell);
    }
    C->map_entries = next;
  }
  C->d_ovmap.release();
  C->d_moff.release();
  if (!jobs.empty()) {
    KAH_TRY(C->upload(C->d_moff, mbase.data(), mbase.size(), "overlap map bases"));
    KAH_TRY(C->alloc(C->d_ovmap, sizeof(uint32_t) * size_t(C->map_entries), "overlap maps"));
    DBuf d_jobs;
    KAH_TRY(C->upload(d_jobs, jobs.data(), jobs.size(), "overlap jobs"));
    DevArgs A{};
    A.lmax = C->lmax;
    A.glev = C->d_glev.as<int8_t>();
    A.gkey = C->d_gkey.as<uint32_t>();
    for (size_t j0 = 0; j0 < jobs.size(); j0 += (1u << 30)) {
      const int nj = int(std::min<size_t>(jobs.size() - j0, size_t(1) << 30));
      KAH_TRY(C->hip(launch_overlap_map(A, D, d_jobs.as<OvJob>() + j0, nj,
                                        C->d_ovmap.as<uint32_t>(), C->stream),
                     "overlap map launch"));
    }
    KAH_TRY(C->hip(hipStreamSynchronize(C->stream), "overlap maps"));
  }
  std::vector<int32_t> hoff(size_t(nc + 1), 0), eoff(size_t(nc + 1), 0);
  std::vector<HalfDesc> halves;
  std::vector<EntryDesc> ents;
  for (int64_t c = 0; c < nc; ++c) {
    hoff[c] = int32_t(halves.size());
    eoff[c] = int32_t(ents.size());
    if (B.hal[c].size() > 24)
      return C->fail(KAH_EUNSUPPORTED, "cell " + std::to_string(c) + " owns more than 24 half fluxes");
    halves.insert(halves.end(), B.hal[c].begin(), B.hal[c].end());
    ents.insert(ents.end(), B.ent[c].begin(), B.ent[c].end());
  }
  hoff[nc] = int32_t(halves.size());
  eoff[nc] = int32_t(ents.size());
  // float copies for the FP32 advection form
  for (auto& sd : B.sl)
    for (int k = 0; k < 2; ++k) {
      sd.rds_f[k] = float(sd.rds[k]);
      sd.inv_cnt_f[k] = float(sd.inv_cnt[k]);
    }
  for (auto& e : B.slnb)
    for (int t = 0; t < 3; ++t) e.dm_f[t] = float(e.dm[t]);
  KAH_TRY(C->upload(C->d_sl, B.sl.data(), B.sl.size(), "slope descriptors"));
  KAH_TRY(C->upload(C->d_slnb, B.slnb.data(), B.slnb.size(), "slope neighbours"));
  KAH_TRY(C->upload(C->d_hoff, hoff.data(), hoff.size(), "half offsets"));
  KAH_TRY(C->upload(C->d_halves, halves.data(), halves.size(), "half descriptors"));
  KAH_TRY(C->upload(C->d_eoff, eoff.data(), eoff.size(), "entry offsets"));
  KAH_TRY(C->upload(C->d_ents, ents.data(), ents.size(), "entry descriptors"));
  KAH_TRY(C->upload(C->d_recds, T->cell_ds, size_t(nc * D), "cell ds"));
  KAH_TRY(C->upload(C->d_cmid, T->cell_mid, size_t(nc * D), "cell midpoints"));
  KAH_TRY(C->upload(C->d_benc, T->bound_enc, size_t(nc), "bound_enc"));
  C->voff_h.assign(T->vs_offset, T->vs_offset + nc + 1);
  std::vector<double> bc(size_t(std::max(T->n_bc, 1) * C->NW), 0.0);
  if (T->n_bc > 0) std::memcpy(bc.data(), T->bc_prim, sizeof(double) * T->n_bc * C->NW);
  KAH_TRY(C->upload(C->d_bc, bc.data(), bc.size(), "bc prim"));

  // fluid cells, per level and per half-count bin
  C->fluid.clear();
  C->phase_cells = 0;
  for (auto* b : C->bins) delete b;
  C->bins.clear();
  int Lmin = 1 << 30;
  for (int64_t c = 0; c < nc; ++c) {
    if (T->bound_enc[c] < 0) continue;
    C->fluid.push_back(int32_t(c));
    C->phase_cells += T->vs_offset[c + 1] - T->vs_offset[c];
    Lmin = std::min(Lmin, int(T->ps_level[c]));
  }
  // Per level and half-flux count: cells owning a wall half need the whole-cell wall density
  // and run the per-cell slope kernel (one partial row per half flux); all others run the
  // chunked kernel (one partial row per chunk of 64*VPT*(3 or 4) velocity cells).
  std::vector<char> walled(size_t(nc), 0);
  for (int64_t c = 0; c < nc; ++c)
    for (int h = hoff[c]; h < hoff[c + 1]; ++h) walled[c] |= (halves[h].mode == MI_WALL);
  for (int32_t c : C->fluid)
    if (eoff[c + 1] - eoff[c] > MAX_ENTRIES)
      return C->fail(KAH_EUNSUPPORTED, "cell " + std::to_string(c) + " has more than " +
                                           std::to_string(MAX_ENTRIES) + " face sides");
  // Chunks of the slope kernel: 3-wave or 4-wave workgroups (64*VPT velocity cells per wave),
  // whichever pads the cell's grid least (a 568-cell grid fills one 3-wave chunk of 576 at
  // VPT 3; ties go to 4 waves).
  auto wpi_of = [&](int64_t c) {
    const int64_t n = T->vs_offset[c + 1] - T->vs_offset[c];
    auto padded = [&](int64_t w) {
      const int64_t cv = w * 64 * C->slope_vpt;
      return (n + cv - 1) / cv * cv;
    };
    return padded(3) < padded(4) ? 3 : 4;
  };
  auto nchunks = [&](int64_t c) {
    const int64_t cv = int64_t(C->slope_vpt) * 64 * wpi_of(c);
    return int((T->vs_offset[c + 1] - T->vs_offset[c] + cv - 1) / cv);
  };
  // Launch order of the slopes: the reference sweeps coarse to fine levels because the
  // transverse projection (update_slope_transverse_level!, src/Flux/Slope.jl:849-945) reads
  // the final slopes of coarser neighbours. Each cell gets the depth of that dependency (0 when
  // its stencil projects no other cell's slopes, else 1 + the depth of those neighbours), and
  // the launches go by depth: every cell still runs after the cells it reads, and the bulk of
  // the mesh (depth 0) is one launch.
  std::vector<int> depth(size_t(nc), 0);
  int Dmax = 0;
  {
    std::vector<int32_t> order(C->fluid);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return T->ps_level[a] < T->ps_level[b]; });
    bool by_level = false;
    for (int32_t c : order) {
      int dc = 0;
      for (int d = 0; d < D; ++d) {
        const SlopeDesc& sd = B.sl[size_t(c) * D + d];
        const int sides = sd.mode == SL_INNER ? 2 : (sd.mode == SL_BOUND ? 1 : 0);
        for (int sx = 0; sx < sides; ++sx)
          for (int e = 0; e < sd.nb_cnt[sx]; ++e) {
            const SlopeNb& nbr = B.slnb[size_t(sd.nb_beg[sx] + e)];
            if (!nbr.proj) continue;
            if (T->ps_level[nbr.cell] >= T->ps_level[c]) by_level = true;
            dc = std::max(dc, 1 + depth[size_t(nbr.cell)]);
          }
      }
      depth[size_t(c)] = dc;
    }
    if (by_level)   // a projection from a level that is not coarser: keep the level order
      for (int32_t c : C->fluid) depth[size_t(c)] = T->ps_level[c] - Lmin;
    for (int32_t c : C->fluid) Dmax = std::max(Dmax, depth[size_t(c)]);
  }
  // Regular cells of the slope kernel: every direction SL_INNER with one neighbour per side,
  // no transverse projection, and 2*D half fluxes that are the cell's own limited or
  // plain-slope reconstructions.
  // Returns 0 (not regular), 1 (both neighbours of every direction on the cell's velocity
  // grid) or 2 (some on other grids).
  auto slope_regular = [&](int32_t c) {
    if (!C->reg_mode || hoff[c + 1] - hoff[c] != 2 * D) return 0;
    int cls = 1;
    for (int d = 0; d < D; ++d) {
      const SlopeDesc& sd = B.sl[size_t(c) * D + d];
      if (sd.mode != SL_INNER || sd.nb_cnt[0] != 1 || sd.nb_cnt[1] != 1) return 0;
      for (int sx = 0; sx < 2; ++sx) {
        const SlopeNb& nbr = B.slnb[size_t(sd.nb_beg[sx])];
        if (nbr.proj) return 0;
        if (!C->uni && nbr.map != MAP_SAME) cls = 2;
      }
    }
    for (int h = hoff[c]; h < hoff[c + 1]; ++h) {
      const HalfDesc& hd = halves[size_t(h)];
      if (hd.data_cell != c || (hd.mode != MI_LIMITED && hd.mode != MI_UNLIMITED)) return 0;
    }
    return cls;
  };
  for (int L = 0; L <= Dmax; ++L) {
    // half-count bins: the per-thread accumulators of K1 are MAXH*(dim+2) doubles, so a bin
    // sized to the common counts (4: 2D interior, 6: 3D interior) keeps registers low
    // bins 5 and 6: regular cells (the slopes and half fluxes of slope_regular /
    // half_regular), with every neighbour on the cell's grid, or some on other grids
    const int caps[7] = {4, 6, 8, 16, 24, -1, -2};
    std::vector<int32_t> lists[7];
    std::vector<int2> works[7][2];   // [bin][3-wave items, 4-wave items]
    for (int32_t c : C->fluid) {
      if (depth[size_t(c)] != L) continue;
      int nh = hoff[c + 1] - hoff[c];
      int b = nh <= 4 ? 0 : (nh <= 6 ? 1 : (nh <= 8 ? 2 : (nh <= 16 ? 3 : 4)));
      if (walled[c]) {
        lists[b].push_back(c);
      } else {
        const int rc = slope_regular(c);
        if (rc) b = 4 + rc;
        auto& wl = works[b][wpi_of(c) == 3 ? 0 : 1];
        for (int ch = 0; ch < nchunks(c); ++ch) wl.push_back(make_int2(c, ch));
      }
    }
    for (int b = 0; b < 7; ++b) {
      for (int kind = 0; kind < 3; ++kind) {   // per-cell kernel, 3-wave items, 4-wave items
        const size_t cnt = kind ? works[b][kind - 1].size() : lists[b].size();
        if (cnt == 0) continue;
        LevelBin* lb = new LevelBin();
        lb->level = L;
        lb->maxh = caps[b];
        lb->n = int(cnt);
        lb->chunked = kind > 0;
        lb->wpi = kind == 1 ? 3 : 4;
        C->bins.push_back(lb);
        if (kind) KAH_TRY(C->upload(lb->items, works[b][kind - 1].data(), cnt, "slope work items"));
        else KAH_TRY(C->upload(lb->items, lists[b].data(), cnt, "level cells"));
      }
    }
  }
  KAH_TRY(C->upload(C->d_fluid, C->fluid.data(), C->fluid.size(), "fluid cells"));

  // partial rows of every half flux
  std::vector<int64_t> poff(size_t(std::max<int64_t>(C->npairs, 1) * 2), 0);
  std::vector<int32_t> pcnt(poff.size(), 0);
  for (int64_t c = 0; c < nc; ++c)
    for (int h = hoff[c]; h < hoff[c + 1]; ++h)
      pcnt[halves[h].slot] = walled[c] ? 1 : nchunks(c);
  for (int64_t p = 0; p < int64_t(C->ib_of_pair.size()); ++p)
    if (C->ib_of_pair[size_t(p)] >= 0) pcnt[size_t(2 * p + 1)] = 1;   // written by the IB kernel
  int64_t rows = 0;
  for (size_t s = 0; s < poff.size(); ++s) {
    poff[s] = rows;
    rows += pcnt[s];
  }
  C->part_rows = rows;
  KAH_TRY(C->upload(C->d_poff, poff.data(), poff.size(), "partial offsets"));
  KAH_TRY(C->upload(C->d_pcnt, pcnt.data(), pcnt.size(), "partial counts"));

  // state buffers
  const int64_t NDF = C->NDF, NW = C->NW;
  const size_t fs = C->fsize();
  // the state buffers hold the SolidNeighbor rows of the IB faces after the cells' rows
  const int64_t prow = C->np + C->ib_rows;
  KAH_TRY(C->alloc(C->d_df[0], fs * prow * NDF, "df"));
  KAH_TRY(C->alloc(C->d_df[1], fs * prow * NDF, "df"));
  if (C->ib_nf > 0) {
    KAH_TRY(C->zero(C->d_df[0]));
    KAH_TRY(C->zero(C->d_df[1]));
  }
  KAH_TRY(C->alloc(C->d_sdf, fs * prow * NDF * D, "sdf"));   // + the SolidNeighbors' slopes
  KAH_TRY(C->alloc(C->d_oflux, fs * prow * NDF * D, "outgoing micro fluxes"));
  KAH_TRY(C->alloc(C->d_psflux, sizeof(double) * nc * NW, "flux"));
  KAH_TRY(C->alloc(C->d_w, sizeof(double) * nc * NW, "w"));
  KAH_TRY(C->alloc(C->d_prim, sizeof(double) * nc * NW, "prim"));
  KAH_TRY(C->alloc(C->d_qf, sizeof(double) * nc * D, "qf"));
  KAH_TRY(C->alloc(C->d_half, sizeof(double) * std::max<int64_t>(C->part_rows, 1) * NW, "half fluxes"));
  KAH_TRY(C->alloc(C->d_rho, sizeof(double) * std::max<int64_t>(C->npairs, 1), "wall density"));
  KAH_TRY(C->alloc(C->d_res, sizeof(double) * nc * 2 * NW, "residual"));
  if (C->P.marching != KAH_MARCH_CIP)
    KAH_TRY(C->alloc(C->d_cpre, sizeof(double) * nc * cpre_len(D), "cell prologue records"));
  else
    C->d_cpre.release();
  C->d_vflux.release();
  // f* and M(prim_c) of the cell in LDS, in the per-velocity arithmetic type (float in the
  // FP32 form, except under CIP_Marching)
  const size_t sc_elem =
      (C->f32 && C->P.marching != KAH_MARCH_CIP) ? sizeof(float) : sizeof(double);
  C->lds_bytes = sc_elem * size_t(NDF + 1) * size_t(C->max_n);
  if (C->lds_bytes > C->lds_limit) C->lds_bytes = 0;
  // Flux/relaxation launches by grid size (adapted grids): grids up to 640, 1280, 2560, ...
  // velocity cells, each launch with the LDS of its largest grid, so that the many small grids
  // of an adapted mesh are not held to the LDS footprint of the largest one; and within a size,
  // regular cells (2*DIM faces to fluid cells on the same velocity grid, limited or plain-slope
  // reconstruction on both sides) apart from the others.
  for (auto* b : C->frbins) delete b;
  C->frbins.clear();
  {
    // the regular gather reads the outgoing micro fluxes the slope kernel stores while it
    // forms the half fluxes (CAIDVM only), so neither side may be a cell whose slope kernel
    // forms them in another arithmetic (the per-cell kernel of walled cells; in the FP32 form
    // CIP_Marching gathers in FP64 while the slope kernel reconstructs in FP32)
    const bool caidvm = C->P.flux_kind == KAH_FLUX_CAIDVM;
    const bool same_arith = !(C->f32 && C->P.marching == KAH_MARCH_CIP);
    // 0: not regular; 1: every neighbour on the cell's velocity grid; 2: some on other grids
    auto regular = [&](int32_t c) {
      if (!C->reg_mode || !caidvm || !same_arith || walled[c] || eoff[c + 1] - eoff[c] != 2 * D)
        return 0;
      int cls = 1;
      for (int e = eoff[c]; e < eoff[c + 1]; ++e) {
        const EntryDesc& ed = ents[size_t(e)];
        if (ed.oth_cell < 0 || ed.oth_cell >= nc || T->bound_enc[ed.oth_cell] < 0) return 0;
        // a neighbour on another level shares a hanging face: its side holds more than one
        // face, each with its own reconstruction, so one stored outgoing flux cannot serve
        if (walled[ed.oth_cell] || T->ps_level[ed.oth_cell] != T->ps_level[c]) return 0;
        if (!C->uni && ed.map != MAP_SAME) cls = 2;
        if ((ed.own_mode != MI_LIMITED && ed.own_mode != MI_UNLIMITED) ||
            (ed.oth_mode != MI_LIMITED && ed.oth_mode != MI_UNLIMITED))
          return 0;
      }
      return cls;
    };
    // keys: 3*size bin + regular class
    std::map<int, std::vector<int32_t>> lists;
    std::map<int, int64_t> bmax;
    for (int32_t c : C->fluid) {
      const int64_t n = T->vs_offset[c + 1] - T->vs_offset[c];
      int b = 0;
      if (!C->uni)
        while ((int64_t(640) << b) < n) ++b;
      const int key = 3 * b + regular(c);
      lists[key].push_back(c);
      bmax[key] = std::max(bmax[key], n);
    }
    bool need_scratch = false;
    for (auto& kv : lists) {
      const int key = kv.first;
      auto* fb = new kah_ctx::FrBin();
      C->frbins.push_back(fb);
      fb->n = int(kv.second.size());
      fb->reg_class = key % 3;
      // LDS and workgroup size from the size class, so that a cell gets the same launch shape
      // (and the same order of its sums) whether it is regular or not
      const int64_t cap = C->uni ? int64_t(C->n0) : std::min<int64_t>(int64_t(640) << (key / 3),
                                                                       int64_t(C->max_n));
      fb->lds = sc_elem * size_t(NDF + 1) * size_t(std::max(cap, bmax[key]));
      if (fb->lds > C->lds_limit) fb->lds = 0;
      fb->waves = C->flux_waves(fb->lds);
      need_scratch |= fb->lds == 0;
      KAH_TRY(C->upload(fb->cells, kv.second.data(), kv.second.size(), "flux cells"));
    }
    if (need_scratch) KAH_TRY(C->alloc(C->d_scratch, sizeof(double) * C->np * (NDF + 1), "scratch"));
    else C->d_scratch.release();
  }
  KAH_TRY(C->zero(C->d_sdf));
  KAH_TRY(C->zero(C->d_psflux));
  KAH_TRY(C->zero(C->d_half));
  KAH_TRY(C->zero(C->d_rho));
  KAH_TRY(C->hip(hipStreamSynchronize(C->stream), "topology upload"));
  C->cur = 0;
  C->have_topo = true;
  return KAH_OK;
}

}  // extern "C"

namespace {

DevArgs make_args(kah_ctx* C, double dt) {
  DevArgs A{};
  A.uni_n = C->n0;
  A.lmax = C->lmax;
  A.vt_base = C->d_vtb.as<int32_t>();
  A.ovmap = C->d_ovmap.as<uint32_t>();
  A.map_base = C->d_moff.as<int64_t>();
  A.gof = C->d_gof.as<int64_t>();
  A.vt_mode = C->vt_mode;
  A.vt_n0 = C->vt_n0;
  A.vt_n1 = C->vt_n1;
  A.voff = C->d_voff.as<int64_t>();
  A.gu = C->d_gu.as<double>();
  A.gw = C->d_gw.as<double>();
  A.glev = C->d_glev.as<int8_t>();
  A.gkey = C->d_gkey.as<uint32_t>();
  A.df = C->d_df[C->cur].as<double>();
  A.df_out = C->d_df[1 - C->cur].as<double>();
  A.sdf = C->d_sdf.as<double>();
  A.oflux = C->d_oflux.as<double>();
  A.vflux = C->d_vflux.as<double>();
  A.psflux = C->d_psflux.as<double>();
  A.w = C->d_w.as<double>();
  A.prim = C->d_prim.as<double>();
  A.qf = C->d_qf.as<double>();
  A.half_fw = C->d_half.as<double>();
  A.part_off = C->d_poff.as<int64_t>();
  A.part_cnt = C->d_pcnt.as<int32_t>();
  A.rho_w = C->d_rho.as<double>();
  A.res = C->d_res.as<double>();
  A.cpre = C->d_cpre.p ? C->d_cpre.as<double>() : nullptr;
  A.scratch = C->d_scratch.as<double>();
  A.sl = C->d_sl.as<SlopeDesc>();
  A.slnb = C->d_slnb.as<SlopeNb>();
  A.half_off = C->d_hoff.as<int32_t>();
  A.halves = C->d_halves.as<HalfDesc>();
  A.ent_off = C->d_eoff.as<int32_t>();
  A.ents = C->d_ents.as<EntryDesc>();
  A.rec_ds = C->d_recds.as<double>();
  A.bc_prim = C->d_bc.as<double>();
  A.sdf_r = A.sdf;
  A.half_fw_r = A.half_fw;
  A.rho_w_r = A.rho_w;
  A.K = C->P.K;
  A.gamma = C->P.gamma;
  A.Pr = C->P.Pr;
  A.mu = C->P.mu_ref;
  A.omega = C->P.omega;
  A.dt = dt;
  A.flux_kind = C->P.flux_kind;
  const bool ppib = C->P.marching == KAH_MARCH_CIP && C->ib_nf > 0;
  A.ib_faces = ppib ? C->d_ib_faces.as<IbFaceDesc>() : nullptr;
  A.ib_ssdf = ppib ? C->d_ib_ssdf.as<double>() : nullptr;
  A.ib_base = int32_t(C->nc);
  return A;
}

int check_ready(kah_ctx* C, double dt) {
  if (!C->have_topo) return C->fail(KAH_ESTATE, "no topology: call kah_set_topology first");
  if (!C->have_state) return C->fail(KAH_ESTATE, "no state: call kah_set_state first");
  if (!(dt > 0.0) || !std::isfinite(dt)) return C->fail(KAH_EINVAL, "dt must be positive and finite");
  C->ev_next = 0;   // the fork/join events are reused call after call
  return C->hip(hipSetDevice(C->P.device), "hipSetDevice");
}

int ev_begin(kah_ctx* C, EvPair& e, int cls) {
  e.cls = cls;
  KAH_TRY(C->hip(hipEventCreate(&e.a), "hipEventCreate"));
  KAH_TRY(C->hip(hipEventCreate(&e.b), "hipEventCreate"));
  return C->hip(hipEventRecord(e.a, C->stream), "hipEventRecord");
}
int ev_end(kah_ctx* C, EvPair& e) {
  KAH_TRY(C->hip(hipEventRecord(e.b, C->stream), "hipEventRecord"));
  C->events.push_back(e);
  return 0;
}

// Event of the fork/join pool (created on first use; every fork and join point of a step has
// its own, so that a captured graph records each dependency once).
int pool_event(kah_ctx* C, hipEvent_t& ev) {
  if (C->ev_next == C->ev_pool.size()) {
    hipEvent_t e = nullptr;
    KAH_TRY(C->hip(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate"));
    C->ev_pool.push_back(e);
  }
  ev = C->ev_pool[C->ev_next++];
  return 0;
}
// Streams of the k-th branch: 0 is the context's stream.
hipStream_t branch_stream(kah_ctx* C, int k) { return k == 0 ? C->stream : C->side[k - 1]; }
int fork_branches(kah_ctx* C, int nb) {
  if (nb <= 1) return 0;
  hipEvent_t ev;
  KAH_TRY(pool_event(C, ev));
  KAH_TRY(C->hip(hipEventRecord(ev, C->stream), "hipEventRecord"));
  for (int k = 1; k < nb; ++k)
    KAH_TRY(C->hip(hipStreamWaitEvent(C->side[k - 1], ev, 0), "hipStreamWaitEvent"));
  return 0;
}
int join_branches(kah_ctx* C, int nb) {
  for (int k = 1; k < nb; ++k) {
    hipEvent_t ev;
    KAH_TRY(pool_event(C, ev));
    KAH_TRY(C->hip(hipEventRecord(ev, C->side[k - 1]), "hipEventRecord"));
    KAH_TRY(C->hip(hipStreamWaitEvent(C->stream, ev, 0), "hipStreamWaitEvent"));
  }
  return 0;
}

int run_slope(kah_ctx* C, const DevArgs& A) {
  EvPair e{};
  if (C->timing) KAH_TRY(ev_begin(C, e, 0));
  // bins of one projection depth are independent; a depth waits for the one before
  for (size_t b0 = 0; b0 < C->bins.size();) {
    size_t b1 = b0;
    while (b1 < C->bins.size() && C->bins[b1]->level == C->bins[b0]->level) ++b1;
    const int nb = std::min<int>(C->branches, int(b1 - b0));
    KAH_TRY(fork_branches(C, nb));
    for (size_t j = b0; j < b1; ++j) {
      const LevelBin* b = C->bins[j];
      hipStream_t st = branch_stream(C, int(j - b0) % nb);
      if (b->chunked)
        KAH_TRY(C->hip(launch_slope_chunk(A, C->D, C->uni, C->f32, b->items.as<int2>(), b->n,
                                          b->maxh, C->slope_vpt, b->wpi, st),
                       "slope kernel launch"));
      else
        KAH_TRY(C->hip(launch_slope(A, C->D, C->uni, C->f32, b->items.as<int32_t>(), b->n,
                                    b->maxh, st),
                       "slope kernel launch"));
    }
    KAH_TRY(join_branches(C, nb));
    b0 = b1;
  }
  if (C->timing) KAH_TRY(ev_end(C, e));
  return 0;
}

// update_solid_cell! and update_solid_neighbor! (src/Flux/Flux.jl:461-482)
int run_ib(kah_ctx* C, const DevArgs& A) {
  if (C->ib_ns == 0 && C->ib_nf == 0) return 0;
  KAH_TRY(C->hip(launch_ib(A, C->D, C->uni, C->f32, C->ib_ns, C->d_ib_solid.as<int32_t>(),
                           C->d_ib_snoff.as<int32_t>(), C->d_ib_snb.as<int32_t>(), C->ib_nf,
                           C->d_ib_faces.as<IbFaceDesc>(), C->d_ib_img.as<int32_t>(),
                           C->d_ib_cut.as<int32_t>(), C->d_ib_cutw.as<double2>(),
                           C->d_cmid.as<double>(), C->d_df[C->cur].p, C->d_df[1 - C->cur].p,
                           C->stream),
                 "immersed-boundary kernel launch"));
  if (C->ib_nf > 0) C->ib_done = true;
  return 0;
}

int run_flux_relax(kah_ctx* C, const DevArgs& A, int mode) {
  EvPair e{};
  if (C->timing) KAH_TRY(ev_begin(C, e, 1));
  int march = C->P.marching == KAH_MARCH_EULER ? MARCH_EULER
               : (C->P.marching == KAH_MARCH_CIP ? MARCH_CIP : MARCH_CAIDVM);
  if (mode == MODE_FUSED && march != MARCH_CIP)
    KAH_TRY(C->hip(launch_cell_pre(A, C->D, C->d_fluid.as<int32_t>(), int(C->fluid.size()),
                                   C->stream),
                   "cell prologue launch"));
  // the bins write disjoint cells and read only what the slope phase wrote
  const int nb = std::min<int>(C->branches, int(C->frbins.size()));
  KAH_TRY(fork_branches(C, nb));
  for (size_t j = 0; j < C->frbins.size(); ++j) {
    const kah_ctx::FrBin* b = C->frbins[j];
    const int m = (mode != MODE_FUSED || b->reg_class == 0) ? mode
                  : (b->reg_class == 1 ? MODE_FUSED_REG : MODE_FUSED_REGX);
    KAH_TRY(C->hip(launch_flux_relax(A, C->D, C->uni, C->f32, b->cells.as<int32_t>(), b->n, m,
                                     march, b->lds, C->fr_waves ? C->fr_waves : b->waves,
                                     branch_stream(C, int(j) % nb)),
                   "flux/relaxation kernel launch"));
  }
  KAH_TRY(join_branches(C, nb));
  if (C->timing) KAH_TRY(ev_end(C, e));
  return 0;
}

int finish_iterate(kah_ctx* C, int64_t step, double* sum_res, double* sum_avg) {
  const int NW = C->NW;
  C->cur = 1 - C->cur;
  bool check = (step % C->P.st_check_interval) == 0;
  if (!check || (!sum_res && !sum_avg)) {
    for (int q = 0; q < NW; ++q) {
      if (sum_res) sum_res[q] = 0.0;
      if (sum_avg) sum_avg[q] = 0.0;
    }
    return 0;
  }
  std::vector<double> r(size_t(C->nc * 2 * NW));
  KAH_TRY(C->hip(hipMemcpyAsync(r.data(), C->d_res.p, r.size() * sizeof(double),
                                hipMemcpyDeviceToHost, C->stream), "residual download"));
  KAH_TRY(C->hip(hipStreamSynchronize(C->stream), "residual download"));
  for (int q = 0; q < NW; ++q) {
    double sr = 0.0, sa = 0.0;
    for (int32_t c : C->fluid) {
      sr += r[size_t(c) * 2 * NW + q];
      sa += r[size_t(c) * 2 * NW + NW + q];
    }
    if (sum_res) sum_res[q] = sr;
    if (sum_avg) sum_avg[q] = sa;
  }
  return 0;
}

}  // namespace

extern "C" {

int kah_set_state(kah_ctx* C, const double* w, const double* prim, const double* df) {
  if (!C) return KAH_EINVAL;
  if (!C->have_topo) return C->fail(KAH_ESTATE, "no topology: call kah_set_topology first");
  if (!w || !prim || !df) return C->fail(KAH_EINVAL, "null state array");
  KAH_TRY(C->hip(hipSetDevice(C->P.device), "hipSetDevice"));
  const int64_t NW = C->NW;
  KAH_TRY(C->hip(hipMemcpyAsync(C->d_w.p, w, sizeof(double) * C->nc * NW, hipMemcpyHostToDevice,
                                C->stream), "w upload"));
  KAH_TRY(C->hip(hipMemcpyAsync(C->d_prim.p, prim, sizeof(double) * C->nc * NW,
                                hipMemcpyHostToDevice, C->stream), "prim upload"));
  KAH_TRY(C->upload_f(C->d_df[C->cur], df, size_t(C->np * C->NDF), "df upload"));
  KAH_TRY(C->zero(C->d_psflux));
  KAH_TRY(C->zero(C->d_vflux));
  KAH_TRY(C->zero(C->d_qf));
  KAH_TRY(C->hip(hipStreamSynchronize(C->stream), "state upload"));
  C->have_state = true;
  return KAH_OK;
}

int kah_get_state(kah_ctx* C, double* w, double* prim, double* qf, double* sw, double* df,
                  double* sdf) {
  if (!C) return KAH_EINVAL;
  if (!C->have_state) return C->fail(KAH_ESTATE, "no state");
  KAH_TRY(C->hip(hipSetDevice(C->P.device), "hipSetDevice"));
  const int64_t NW = C->NW, D = C->D;
  if (sw) {
    if (!C->d_sw.p) KAH_TRY(C->alloc(C->d_sw, sizeof(double) * C->nc * NW * D, "sw"));
    KAH_TRY(C->zero(C->d_sw));
    DevArgs A = make_args(C, 0.0);
    KAH_TRY(C->hip(launch_macro_slope(A, C->D, C->uni, C->f32, C->d_fluid.as<int32_t>(),
                                      int(C->fluid.size()), C->d_sw.as<double>(), C->stream),
                   "macro slope launch"));
    KAH_TRY(C->hip(hipMemcpyAsync(sw, C->d_sw.p, sizeof(double) * C->nc * NW * D,
                                  hipMemcpyDeviceToHost, C->stream), "sw download"));
  }
  auto dl = [&](double* dst, const DBuf& b, size_t count, const char* what) -> int {
    if (!dst) return 0;
    return C->hip(hipMemcpyAsync(dst, b.p, count * sizeof(double), hipMemcpyDeviceToHost, C->stream), what);
  };
  KAH_TRY(dl(w, C->d_w, size_t(C->nc * NW), "w download"));
  KAH_TRY(dl(prim, C->d_prim, size_t(C->nc * NW), "prim download"));
  KAH_TRY(dl(qf, C->d_qf, size_t(C->nc * D), "qf download"));
  KAH_TRY(C->download_f(df, C->d_df[C->cur], size_t(C->np * C->NDF), "df download"));
  KAH_TRY(C->download_f(sdf, C->d_sdf, size_t(C->np * C->NDF * D), "sdf download"));
  return C->hip(hipStreamSynchronize(C->stream), "state download");
}

int kah_mirror_layout(kah_ctx* C, const int32_t* cells, int64_t n, int kind, int64_t* offsets) {
  if (!C) return KAH_EINVAL;
  if (!C->have_topo) return C->fail(KAH_ESTATE, "no topology: call kah_set_topology first");
  if (kind != KAH_MIRROR_DATA && kind != KAH_MIRROR_SLOPE)
    return C->fail(KAH_EINVAL, "unknown mirror buffer kind " + std::to_string(kind));
  if (n < 0 || n >= (int64_t(1) << 31)) return C->fail(KAH_EINVAL, "mirror count out of range");
  if (!offsets || (n > 0 && !cells)) return C->fail(KAH_EINVAL, "null mirror array");
  const int64_t D = C->D, NW = C->NW, NDF = C->NDF;
  offsets[0] = 0;
  for (int64_t k = 0; k < n; ++k) {
    const int32_t c = cells[k];
    if (c < 0 || c >= C->nc) return C->fail(KAH_EINVAL, "mirror cell id out of range");
    const int64_t nv = C->benc[c] < 0 ? 0 : C->voff_h[c + 1] - C->voff_h[c];
    // size_Ghost_Data / size_Ghost_Slope (src/Parallel/Ghost.jl:1-2)
    const int64_t len = kind == KAH_MIRROR_DATA ? 3 * D + 4 + NDF * nv
                                                 : std::max<int64_t>(nv * NDF * D + D * NW, 1);
    offsets[k + 1] = offsets[k] + len;
  }
  return KAH_OK;
}

}  // extern "C"

namespace {

// shared part of kah_pack_mirrors / kah_unpack_mirrors: offsets, cell list and staging buffer
// on the device
int mirror_prepare(kah_ctx* C, const int32_t* cells, int64_t n, int kind,
                   std::vector<int64_t>& off) {
  if (!C->have_state) return C->fail(KAH_ESTATE, "no state: call kah_set_state first");
  off.resize(size_t(n) + 1);
  KAH_TRY(kah_mirror_layout(C, cells, n, kind, off.data()));
  KAH_TRY(C->hip(hipSetDevice(C->P.device), "hipSetDevice"));
  if (n == 0) return KAH_OK;
  KAH_TRY(C->reserve(C->d_xcells, sizeof(int32_t) * size_t(n), "mirror cells"));
  KAH_TRY(C->reserve(C->d_xoff, sizeof(int64_t) * size_t(n), "mirror offsets"));
  KAH_TRY(C->reserve(C->d_xbuf, sizeof(double) * size_t(off[n]), "mirror buffer"));
  KAH_TRY(C->hip(hipMemcpyAsync(C->d_xcells.p, cells, sizeof(int32_t) * size_t(n),
                                hipMemcpyHostToDevice, C->stream), "mirror cells upload"));
  KAH_TRY(C->hip(hipMemcpyAsync(C->d_xoff.p, off.data(), sizeof(int64_t) * size_t(n),
                                hipMemcpyHostToDevice, C->stream), "mirror offsets upload"));
  if (!C->d_sw.p) KAH_TRY(C->alloc(C->d_sw, sizeof(double) * C->nc * C->NW * C->D, "sw"));
  return KAH_OK;
}

int mirror_launch(kah_ctx* C, int64_t n, int kind, int unpack) {
  DevArgs A = make_args(C, 0.0);
  return C->hip(launch_mirror(A, C->D, C->uni, C->f32, C->d_xcells.as<int32_t>(),
                              C->d_xoff.as<int64_t>(), int(n), kind, unpack,
                              C->d_xbuf.as<double>(), C->d_cmid.as<double>(),
                              C->d_benc.as<int32_t>(), C->d_sw.as<double>(),
                              C->d_df[C->cur].p, C->stream),
                "mirror kernel launch");
}

}  // namespace

extern "C" {

int kah_pack_mirrors(kah_ctx* C, const int32_t* cells, int64_t n, int kind, double* out) {
  if (!C) return KAH_EINVAL;
  std::vector<int64_t> off;
  KAH_TRY(mirror_prepare(C, cells, n, kind, off));
  if (n == 0) return KAH_OK;
  if (!out) return C->fail(KAH_EINVAL, "null mirror buffer");
  if (kind == KAH_MIRROR_SLOPE) {
    // PsData.sw of the listed fluid cells, the w-moments of sdf (src/Flux/Slope.jl:1022-1036)
    std::vector<int32_t> fl;
    for (int64_t k = 0; k < n; ++k)
      if (C->benc[cells[k]] >= 0) fl.push_back(cells[k]);
    if (!fl.empty()) {
      KAH_TRY(C->reserve(C->d_xfluid, sizeof(int32_t) * fl.size(), "mirror fluid cells"));
      KAH_TRY(C->hip(hipMemcpyAsync(C->d_xfluid.p, fl.data(), sizeof(int32_t) * fl.size(),
                                    hipMemcpyHostToDevice, C->stream), "mirror cells upload"));
      DevArgs A = make_args(C, 0.0);
      KAH_TRY(C->hip(launch_macro_slope(A, C->D, C->uni, C->f32, C->d_xfluid.as<int32_t>(),
                                        int(fl.size()), C->d_sw.as<double>(), C->stream),
                     "macro slope launch"));
    }
  }
  KAH_TRY(mirror_launch(C, n, kind, 0));
  KAH_TRY(C->hip(hipMemcpyAsync(out, C->d_xbuf.p, sizeof(double) * size_t(off[n]),
                                hipMemcpyDeviceToHost, C->stream), "mirror buffer download"));
  return C->hip(hipStreamSynchronize(C->stream), "mirror pack");
}

int kah_unpack_mirrors(kah_ctx* C, const int32_t* cells, int64_t n, int kind, const double* in) {
  if (!C) return KAH_EINVAL;
  std::vector<int64_t> off;
  KAH_TRY(mirror_prepare(C, cells, n, kind, off));
  if (n == 0) return KAH_OK;
  if (!in) return C->fail(KAH_EINVAL, "null mirror buffer");
  if (kind == KAH_MIRROR_DATA) {
    // the receiving side tells a solid ghost by w[1] == Inf (src/Parallel/Ghost.jl:725-728)
    const int64_t D = C->D;
    for (int64_t k = 0; k < n; ++k) {
      const double* b = in + off[k];
      const bool solid = C->benc[cells[k]] < 0;
      if ((b[2 * D] == std::numeric_limits<double>::infinity()) != solid)
        return C->fail(KAH_EINVAL, "mirror buffer " + std::to_string(k) +
                                       (solid ? " holds fluid data for a solid cell"
                                              : " holds solid data for a fluid cell"));
      const int64_t nv = C->voff_h[cells[k] + 1] - C->voff_h[cells[k]];
      if (!solid && b[3 * D + 2] != double(nv))
        return C->fail(KAH_EINVAL, "mirror buffer " + std::to_string(k) + ": vs_num " +
                                       std::to_string(b[3 * D + 2]) + " != " +
                                       std::to_string(nv) + " velocity cells of cell " +
                                       std::to_string(cells[k]));
    }
  }
  KAH_TRY(C->hip(hipMemcpyAsync(C->d_xbuf.p, in, sizeof(double) * size_t(off[n]),
                                hipMemcpyHostToDevice, C->stream), "mirror buffer upload"));
  KAH_TRY(mirror_launch(C, n, kind, 1));
  return C->hip(hipStreamSynchronize(C->stream), "mirror unpack");
}

int kah_restart_sizes(kah_ctx* C, int64_t* n_cells, int64_t* n_vs) {
  if (!C) return KAH_EINVAL;
  if (!C->have_topo) return C->fail(KAH_ESTATE, "no topology: call kah_set_topology first");
  int64_t nv = 0;
  for (int64_t c = 0; c < C->nc; ++c)
    if (C->benc[c] >= 0) nv += C->voff_h[c + 1] - C->voff_h[c];
  if (n_cells) *n_cells = C->nc;
  if (n_vs) *n_vs = nv;
  return KAH_OK;
}

}  // extern "C"

namespace {

// first restart row of each cell (-1 inside a solid) and the row count
int64_t restart_rows(const kah_ctx* C, std::vector<int64_t>& row) {
  row.resize(size_t(C->nc));
  int64_t r = 0;
  for (int64_t c = 0; c < C->nc; ++c) {
    if (C->benc[c] < 0) {
      row[c] = -1;
      continue;
    }
    row[c] = r;
    r += C->voff_h[c + 1] - C->voff_h[c];
  }
  return r;
}

}  // namespace

extern "C" {

int kah_save_restart(kah_ctx* C, int64_t* vs_nums, int64_t* bound_encs, double* ws,
                     int8_t* vs_levels, double* vs_midpoints, double* vs_df) {
  if (!C) return KAH_EINVAL;
  if (!C->have_state) return C->fail(KAH_ESTATE, "no state: call kah_set_state first");
  if (!vs_nums || !bound_encs || !ws || !vs_levels || !vs_midpoints || !vs_df)
    return C->fail(KAH_EINVAL, "null restart array");
  KAH_TRY(C->hip(hipSetDevice(C->P.device), "hipSetDevice"));
  const int64_t N = C->nc, NW = C->NW, D = C->D, NDF = C->NDF;
  std::vector<int64_t> row;
  const int64_t Nv = restart_rows(C, row);
  for (int64_t c = 0; c < N; ++c) {
    vs_nums[c] = C->benc[c] < 0 ? 0 : C->voff_h[c + 1] - C->voff_h[c];
    bound_encs[c] = C->benc[c];
  }
  DBuf d_row, d_ws, d_lev, d_mid, d_df;   // staging, freed on return
  KAH_TRY(C->upload(d_row, row.data(), row.size(), "restart rows"));
  KAH_TRY(C->alloc(d_ws, sizeof(double) * size_t(N * NW), "restart ws"));
  KAH_TRY(C->alloc(d_lev, size_t(Nv), "restart vs_levels"));
  KAH_TRY(C->alloc(d_mid, sizeof(double) * size_t(Nv * D), "restart vs_midpoints"));
  KAH_TRY(C->alloc(d_df, sizeof(double) * size_t(Nv * NDF), "restart vs_df"));
  KAH_TRY(C->zero(d_ws));   // cells inside a solid keep w = 0 (src/IO/Restart.jl:83)
  DevArgs A = make_args(C, 0.0);
  KAH_TRY(C->hip(launch_restart(A, C->D, C->uni, C->f32, d_row.as<int64_t>(), N, Nv, 0,
                                d_ws.as<double>(), d_lev.as<int8_t>(), d_mid.as<double>(),
                                d_df.as<double>(), C->d_df[C->cur].p, C->stream),
                 "restart kernel launch"));
  KAH_TRY(C->hip(hipMemcpyAsync(ws, d_ws.p, d_ws.bytes, hipMemcpyDeviceToHost, C->stream),
                 "restart ws download"));
  if (Nv > 0) {
    KAH_TRY(C->hip(hipMemcpyAsync(vs_levels, d_lev.p, size_t(Nv), hipMemcpyDeviceToHost,
                                  C->stream), "restart vs_levels download"));
    KAH_TRY(C->hip(hipMemcpyAsync(vs_midpoints, d_mid.p, sizeof(double) * size_t(Nv * D),
                                  hipMemcpyDeviceToHost, C->stream),
                   "restart vs_midpoints download"));
    KAH_TRY(C->hip(hipMemcpyAsync(vs_df, d_df.p, sizeof(double) * size_t(Nv * NDF),
                                  hipMemcpyDeviceToHost, C->stream), "restart vs_df download"));
  }
  return C->hip(hipStreamSynchronize(C->stream), "restart save");
}

int kah_load_restart(kah_ctx* C, const int64_t* vs_nums, const int64_t* bound_encs,
                     const double* ws, const double* vs_df) {
  if (!C) return KAH_EINVAL;
  if (!C->have_topo) return C->fail(KAH_ESTATE, "no topology: call kah_set_topology first");
  if (!vs_nums || !bound_encs || !ws || !vs_df) return C->fail(KAH_EINVAL, "null restart array");
  KAH_TRY(C->hip(hipSetDevice(C->P.device), "hipSetDevice"));
  const int64_t N = C->nc, NW = C->NW, NDF = C->NDF;
  for (int64_t c = 0; c < N; ++c) {
    const int64_t nv = C->benc[c] < 0 ? 0 : C->voff_h[c + 1] - C->voff_h[c];
    if (vs_nums[c] != nv || bound_encs[c] != C->benc[c])
      return C->fail(KAH_EINVAL, "restart arrays do not match the topology at cell " +
                                     std::to_string(c) + " (vs_num " +
                                     std::to_string(vs_nums[c]) + ", bound_enc " +
                                     std::to_string(bound_encs[c]) + ")");
  }
  std::vector<int64_t> row;
  const int64_t Nv = restart_rows(C, row);
  DBuf d_row, d_ws, d_df;
  KAH_TRY(C->upload(d_row, row.data(), row.size(), "restart rows"));
  KAH_TRY(C->upload(d_ws, ws, size_t(N * NW), "restart ws"));
  KAH_TRY(C->upload(d_df, vs_df, size_t(Nv * NDF), "restart vs_df"));
  DevArgs A = make_args(C, 0.0);
  KAH_TRY(C->hip(launch_restart(A, C->D, C->uni, C->f32, d_row.as<int64_t>(), N, Nv, 1,
                                d_ws.as<double>(), nullptr, nullptr, d_df.as<double>(),
                                C->d_df[C->cur].p, C->stream),
                 "restart kernel launch"));
  // the loaded state starts with zero slopes and flux accumulators (src/IO/Restart.jl:317-318)
  KAH_TRY(C->zero(C->d_sdf));
  KAH_TRY(C->zero(C->d_psflux));
  KAH_TRY(C->zero(C->d_vflux));
  KAH_TRY(C->zero(C->d_qf));
  KAH_TRY(C->hip(hipStreamSynchronize(C->stream), "restart load"));
  C->have_state = true;
  return KAH_OK;
}

int kah_vs_resolution(kah_ctx* C, double* vr) {
  if (!C) return KAH_EINVAL;
  if (!C->have_state) return C->fail(KAH_ESTATE, "no state: call kah_set_state first");
  if (!vr) return C->fail(KAH_EINVAL, "null vr");
  if (!C->vs_ok) return C->fail(KAH_EINVAL, "velocity root grid (vs_trees_num / vs_lo / vs_hi) not set");
  KAH_TRY(C->hip(hipSetDevice(C->P.device), "hipSetDevice"));
  DBuf d_res;
  KAH_TRY(C->alloc(d_res, sizeof(double) * size_t(2 * C->nc), "vs resolution"));
  DevArgs A = make_args(C, 0.0);
  KAH_TRY(C->hip(launch_vs_res(A, C->D, C->uni, C->f32, C->vsa, int(C->nc), d_res.as<double>(),
                               C->stream), "vs resolution launch"));
  std::vector<double> r(size_t(2 * C->nc));
  KAH_TRY(C->hip(hipMemcpyAsync(r.data(), d_res.p, d_res.bytes, hipMemcpyDeviceToHost, C->stream),
                 "vs resolution download"));
  KAH_TRY(C->hip(hipStreamSynchronize(C->stream), "vs resolution"));
  // maxima over the fluid cells (vs_resolution(trees), src/Velocity_space/AMR.jl:137-151)
  double dres = 0.0, eres = 0.0;
  auto jmax = [](double a, double b) { return (a != a) ? a : (b != b) ? b : (b > a ? b : a); };
  for (int64_t c = 0; c < C->nc; ++c) {
    if (C->benc[c] < 0) continue;
    dres = jmax(r[size_t(2 * c)], dres);
    eres = jmax(r[size_t(2 * c + 1)], eres);
  }
  vr[0] = dres;
  vr[1] = eres;
  return KAH_OK;
}

int kah_vs_resolution_noalloc(kah_ctx* C, double* vr) {
  if (!C) return KAH_EINVAL;
  if (!C->have_state) return C->fail(KAH_ESTATE, "no state: call kah_set_state first");
  if (!vr) return C->fail(KAH_EINVAL, "null vr");
  if (!C->vs_ok) return C->fail(KAH_EINVAL, "velocity root grid (vs_trees_num / vs_lo / vs_hi) not set");
  if (C->nc > NOALLOC_MAX_CELLS)
    return C->fail(KAH_EUNSUPPORTED, "kah_vs_resolution_noalloc: " + std::to_string(C->nc) +
                                         " cells exceed its scratch limit of " +
                                         std::to_string(NOALLOC_MAX_CELLS) + " cells");
  const int dev = C->P.device;
  double* scr = (dev >= 0 && dev < NOALLOC_MAX_DEVICES) ? g_scratch_ptr[dev] : nullptr;
  if (!scr) {
    const hipError_t e = (dev >= 0 && dev < NOALLOC_MAX_DEVICES) ? g_scratch_err[dev]
                                                                 : hipErrorInvalidDevice;
    return C->fail(KAH_EDEVICE, std::string("kah_vs_resolution_noalloc: no scratch on device ") +
                                    std::to_string(dev) + ": " + hipGetErrorString(e));
  }
  KAH_TRY(C->hip(hipSetDevice(dev), "hipSetDevice"));
  double* res = scr + 2 * NOALLOC_MAX_CELLS;
  DevArgs A = make_args(C, 0.0);
  KAH_TRY(C->hip(launch_vs_res(A, C->D, C->uni, C->f32, C->vsa, int(C->nc), scr, C->stream),
                 "vs resolution launch"));
  KAH_TRY(C->hip(launch_vs_res_max(scr, C->d_benc.as<int32_t>(), int(C->nc), res, C->stream),
                 "vs resolution maxima launch"));
  KAH_TRY(C->hip(hipMemcpyAsync(vr, res, 2 * sizeof(double), hipMemcpyDeviceToHost, C->stream),
                 "vs resolution download"));
  return C->hip(hipStreamSynchronize(C->stream), "vs resolution");
}

int kah_vs_flags(kah_ctx* C, const kah_vs_adapt* p, const double* vr, uint8_t* flags) {
  if (!C) return KAH_EINVAL;
  if (!C->have_state) return C->fail(KAH_ESTATE, "no state: call kah_set_state first");
  if (!p || !vr || !flags) return C->fail(KAH_EINVAL, "null argument");
  if (p->mode != KAH_VS_LOHNER && p->mode != KAH_VS_CONTRIBUTION)
    return C->fail(KAH_EINVAL, "unknown velocity adaptation mode " + std::to_string(p->mode));
  if (!C->vs_ok) return C->fail(KAH_EINVAL, "velocity root grid (vs_trees_num / vs_lo / vs_hi) not set");
  KAH_TRY(C->hip(hipSetDevice(C->P.device), "hipSetDevice"));
  VsArgs P = C->vsa;
  P.mode = p->mode;
  P.cl = p->coeff_local;
  P.cg = p->coeff_global;
  P.tl = p->coeff_lohner;
  P.vr[0] = vr[0];
  P.vr[1] = vr[1];
  DBuf d_fl;
  KAH_TRY(C->alloc(d_fl, size_t(C->np), "vs flags"));
  DevArgs A = make_args(C, 0.0);
  KAH_TRY(C->hip(launch_vs_flags(A, C->D, C->uni, C->f32, P, int(C->nc), d_fl.as<uint8_t>(),
                                 C->stream), "vs flags launch"));
  KAH_TRY(C->hip(hipMemcpyAsync(flags, d_fl.p, size_t(C->np), hipMemcpyDeviceToHost, C->stream),
                 "vs flags download"));
  return C->hip(hipStreamSynchronize(C->stream), "vs flags");
}

int kah_vs_remap(kah_ctx* C, const kah_topology* T, const int32_t* src, const uint8_t* changed) {
  if (!C) return KAH_EINVAL;
  if (!C->have_state) return C->fail(KAH_ESTATE, "no state: call kah_set_state first");
  if (!T || !src || !changed) return C->fail(KAH_EINVAL, "null argument");
  if (T->n_cells != C->nc) return C->fail(KAH_EINVAL, "the new topology has another cell count");
  if (!T->vs_offset) return C->fail(KAH_EINVAL, "null array in topology");
  // every source stays inside its cell's old grid
  const int64_t nc = C->nc, nch = int64_t(1) << C->D;
  for (int64_t c = 0; c < nc; ++c) {
    const int64_t n_old = C->voff_h[c + 1] - C->voff_h[c];
    const int64_t b = T->vs_offset[c], e = T->vs_offset[c + 1];
    if (e <= b) return C->fail(KAH_EINVAL, "empty or unordered velocity grid of cell " + std::to_string(c));
    if (!changed[c] && e - b != n_old)
      return C->fail(KAH_EINVAL, "cell " + std::to_string(c) + " is marked unchanged but its grid size changed");
    for (int64_t i = b; i < e; ++i) {
      const int64_t j = src[i];
      const bool ok = j >= 0 ? j < n_old : (-j - 1) + nch <= n_old;
      if (!ok)
        return C->fail(KAH_EINVAL, "remap source out of range in cell " + std::to_string(c));
      if (!changed[c] && j != i - b)
        return C->fail(KAH_EINVAL, "cell " + std::to_string(c) + " is marked unchanged but remapped");
    }
  }
  KAH_TRY(C->hip(hipSetDevice(C->P.device), "hipSetDevice"));
  KAH_TRY(C->hip(hipStreamSynchronize(C->stream), "remap"));
  // the old state leaves the context; kah_set_topology allocates the new buffers
  DBuf o_df, o_sdf, o_voff, o_w, o_prim, o_qf;
  std::swap(o_df, C->d_df[C->cur]);
  std::swap(o_sdf, C->d_sdf);
  std::swap(o_voff, C->d_voff);
  std::swap(o_w, C->d_w);
  std::swap(o_prim, C->d_prim);
  std::swap(o_qf, C->d_qf);
  KAH_TRY(kah_set_topology(C, T));
  const int64_t NW = C->NW;
  KAH_TRY(C->hip(hipMemcpyAsync(C->d_w.p, o_w.p, sizeof(double) * nc * NW, hipMemcpyDeviceToDevice,
                                C->stream), "w copy"));
  KAH_TRY(C->hip(hipMemcpyAsync(C->d_prim.p, o_prim.p, sizeof(double) * nc * NW,
                                hipMemcpyDeviceToDevice, C->stream), "prim copy"));
  KAH_TRY(C->hip(hipMemcpyAsync(C->d_qf.p, o_qf.p, sizeof(double) * nc * C->D,
                                hipMemcpyDeviceToDevice, C->stream), "qf copy"));
  DBuf d_src, d_ch;
  KAH_TRY(C->upload(d_src, src, size_t(C->np), "remap sources"));
  KAH_TRY(C->upload(d_ch, changed, size_t(nc), "remap changed flags"));
  DevArgs A = make_args(C, 0.0);
  A.df_out = C->d_df[C->cur].as<double>();   // the remapped state is the current f
  KAH_TRY(C->hip(launch_vs_remap(A, C->D, C->f32, o_voff.as<int64_t>(), o_df.p, o_sdf.p,
                                 d_src.as<int32_t>(), d_ch.as<uint8_t>(), int(nc), C->stream),
                 "remap launch"));
  KAH_TRY(C->zero(C->d_psflux));
  KAH_TRY(C->zero(C->d_vflux));
  KAH_TRY(C->hip(hipStreamSynchronize(C->stream), "remap"));
  C->have_state = true;
  return KAH_OK;
}

int kah_vs_project(kah_ctx* C, const int32_t* cells, int64_t n) {
  if (!C) return KAH_EINVAL;
  if (!C->have_state) return C->fail(KAH_ESTATE, "no state: call kah_set_state first");
  if (n < 0 || (n > 0 && !cells)) return C->fail(KAH_EINVAL, "null or negative cell list");
  std::vector<int32_t> list;
  for (int64_t k = 0; k < n; ++k) {
    if (cells[k] < 0 || cells[k] >= C->nc) return C->fail(KAH_EINVAL, "cell id out of range");
    if (C->benc[cells[k]] >= 0) list.push_back(cells[k]);   // solid cells are skipped
  }
  if (list.empty()) return KAH_OK;
  KAH_TRY(C->hip(hipSetDevice(C->P.device), "hipSetDevice"));
  DBuf d_cells, d_scr;
  KAH_TRY(C->upload(d_cells, list.data(), list.size(), "projection cells"));
  const size_t lds = sizeof(double) * size_t(C->NDF + 1) * size_t(C->max_n);
  double* scr = nullptr;
  if (lds > C->lds_limit) {
    KAH_TRY(C->alloc(d_scr, sizeof(double) * size_t(C->np) * size_t(C->NDF + 1), "projection scratch"));
    scr = d_scr.as<double>();
  }
  DevArgs A = make_args(C, 0.0);
  KAH_TRY(C->hip(launch_vs_project(A, C->D, C->uni, C->f32, d_cells.as<int32_t>(),
                                   int(list.size()), lds, scr, C->d_df[C->cur].p, C->stream),
                 "projection launch"));
  return C->hip(hipStreamSynchronize(C->stream), "projection");
}

int kah_get_flux(kah_ctx* C, double* flux, double* vs_flux) {
  if (!C) return KAH_EINVAL;
  if (!C->have_state) return C->fail(KAH_ESTATE, "no state");
  KAH_TRY(C->hip(hipSetDevice(C->P.device), "hipSetDevice"));
  if (flux)
    KAH_TRY(C->hip(hipMemcpyAsync(flux, C->d_psflux.p, sizeof(double) * C->nc * C->NW,
                                  hipMemcpyDeviceToHost, C->stream), "flux download"));
  if (vs_flux) {
    if (C->d_vflux.p)
      KAH_TRY(C->hip(hipMemcpyAsync(vs_flux, C->d_vflux.p, sizeof(double) * C->np * C->NDF,
                                    hipMemcpyDeviceToHost, C->stream), "vs flux download"));
    else
      std::memset(vs_flux, 0, sizeof(double) * C->np * C->NDF);
  }
  return C->hip(hipStreamSynchronize(C->stream), "flux download");
}

int kah_slope(kah_ctx* C, double dt) {
  if (!C) return KAH_EINVAL;
  KAH_TRY(check_ready(C, dt));
  DevArgs A = make_args(C, dt);
  KAH_TRY(run_slope(C, A));
  return C->hip(hipStreamSynchronize(C->stream), "slope");
}

int kah_flux(kah_ctx* C, double dt) {
  if (!C) return KAH_EINVAL;
  KAH_TRY(check_ready(C, dt));
  if (!C->d_vflux.p) {
    KAH_TRY(C->alloc(C->d_vflux, sizeof(double) * C->np * C->NDF, "vs flux"));
    KAH_TRY(C->zero(C->d_vflux));
  }
  DevArgs A = make_args(C, dt);
  KAH_TRY(run_ib(C, A));
  KAH_TRY(run_flux_relax(C, A, MODE_FLUX));
  return C->hip(hipStreamSynchronize(C->stream), "flux");
}

int kah_get_ib(kah_ctx* C, double* aux_df) {
  if (!C) return KAH_EINVAL;
  if (!aux_df) return C->fail(KAH_EINVAL, "null output");
  if (C->ib_nf == 0 || !C->ib_done)
    return C->fail(KAH_ESTATE, "no immersed-boundary faces computed yet");
  KAH_TRY(C->hip(hipSetDevice(C->P.device), "hipSetDevice"));
  // both state buffers hold the rows of the last flux
  const size_t off = size_t(C->np) * C->NDF * C->fsize();
  const size_t cnt = size_t(C->ib_rows) * C->NDF;
  KAH_TRY(C->download_f(aux_df, static_cast<const char*>(C->d_df[C->cur].p) + off, cnt,
                        "ib download"));
  return C->hip(hipStreamSynchronize(C->stream), "ib download");
}

int kah_iterate(kah_ctx* C, double dt, int64_t step, double* sum_res, double* sum_avg) {
  if (!C) return KAH_EINVAL;
  KAH_TRY(check_ready(C, dt));
  if (!C->d_vflux.p) {
    KAH_TRY(C->alloc(C->d_vflux, sizeof(double) * C->np * C->NDF, "vs flux"));
    KAH_TRY(C->zero(C->d_vflux));
  }
  DevArgs A = make_args(C, dt);
  KAH_TRY(run_flux_relax(C, A, MODE_ITER));
  KAH_TRY(C->hip(hipStreamSynchronize(C->stream), "iterate"));
  return finish_iterate(C, step, sum_res, sum_avg);
}

int kah_step(kah_ctx* C, double dt, int64_t step, double* sum_res, double* sum_avg) {
  if (!C) return KAH_EINVAL;
  KAH_TRY(check_ready(C, dt));
  DevArgs A = make_args(C, dt);
  if (C->use_graph && !C->timing) {
    // the launches of one fused step, captured once per buffer parity and dt, then replayed
    // as a single graph launch (the step records no memset or host work)
    const int k = C->cur;
    if (!C->gexec[k] || C->gdt[k] != dt) {
      if (C->gexec[k]) (void)hipGraphExecDestroy(C->gexec[k]);
      C->gexec[k] = nullptr;
      hipGraph_t g = nullptr;
      KAH_TRY(C->hip(hipStreamBeginCapture(C->stream, hipStreamCaptureModeThreadLocal),
                     "hipStreamBeginCapture"));
      int rc = run_slope(C, A);
      if (!rc) rc = run_ib(C, A);
      if (!rc) rc = run_flux_relax(C, A, MODE_FUSED);
      hipError_t ec = hipStreamEndCapture(C->stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      KAH_TRY(C->hip(ec, "hipStreamEndCapture"));
      hipError_t ei = hipGraphInstantiate(&C->gexec[k], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      KAH_TRY(C->hip(ei, "hipGraphInstantiate"));
      C->gdt[k] = dt;
    }
    KAH_TRY(C->hip(hipGraphLaunch(C->gexec[k], C->stream), "hipGraphLaunch"));
  } else {
    KAH_TRY(run_slope(C, A));
    KAH_TRY(run_ib(C, A));
    KAH_TRY(run_flux_relax(C, A, MODE_FUSED));
  }
  bool check = (step % C->P.st_check_interval) == 0 && (sum_res || sum_avg);
  if (check) KAH_TRY(C->hip(hipStreamSynchronize(C->stream), "step"));
  else KAH_TRY(C->hip(hipGetLastError(), "step"));
  return finish_iterate(C, step, sum_res, sum_avg);
}

int kah_sync(kah_ctx* C) {
  if (!C) return KAH_EINVAL;
  return C->hip(hipStreamSynchronize(C->stream), "sync");
}

int kah_copy_bandwidth(kah_ctx* C, int64_t bytes, int reps, double* ms_per_copy) {
  if (!C || bytes <= 0 || reps <= 0) return KAH_EINVAL;
  KAH_TRY(C->hip(hipSetDevice(C->P.device), "hipSetDevice"));
  bytes = (bytes / 16) * 16;
  DBuf a, b;
  KAH_TRY(C->alloc(a, size_t(bytes), "copy src"));
  KAH_TRY(C->alloc(b, size_t(bytes), "copy dst"));
  KAH_TRY(C->zero(a));
  KAH_TRY(C->hip(launch_copy(a.p, b.p, bytes, C->stream), "copy launch"));
  hipEvent_t e0, e1;
  KAH_TRY(C->hip(hipEventCreate(&e0), "hipEventCreate"));
  KAH_TRY(C->hip(hipEventCreate(&e1), "hipEventCreate"));
  KAH_TRY(C->hip(hipEventRecord(e0, C->stream), "hipEventRecord"));
  for (int r = 0; r < reps; ++r)
    KAH_TRY(C->hip(launch_copy(a.p, b.p, bytes, C->stream), "copy launch"));
  KAH_TRY(C->hip(hipEventRecord(e1, C->stream), "hipEventRecord"));
  KAH_TRY(C->hip(hipEventSynchronize(e1), "hipEventSynchronize"));
  float ms = 0.f;
  KAH_TRY(C->hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime"));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms_per_copy) *ms_per_copy = double(ms) / reps;
  return KAH_OK;
}

int64_t kah_phase_cells(const kah_ctx* C) { return C ? C->phase_cells : 0; }
int64_t kah_device_bytes(const kah_ctx* C) { return C ? C->device_bytes() : 0; }

int kah_set_option(kah_ctx* C, int option, int64_t value) {
  if (!C) return KAH_EINVAL;
  switch (option) {
    case KAH_OPT_FLUX_WAVES:
      if (value != 0 && value != 2 && value != 3 && value != 4 && value != 8)
        return C->fail(KAH_EINVAL, "flux waves per workgroup must be 2, 3, 4 or 8 (0: by shape)");
      C->fr_waves = int(value);
      C->drop_graphs();   // captured graphs hold the launch shape
      return KAH_OK;
    case KAH_OPT_GRAPHS:
      if (value != 0 && value != 1) return C->fail(KAH_EINVAL, "graphs must be 0 or 1");
      C->use_graph = value != 0;
      C->drop_graphs();
      return KAH_OK;
    case KAH_OPT_OVERLAP_MAPS:
      if (value != 0 && value != 1) return C->fail(KAH_EINVAL, "overlap maps must be 0 or 1");
      C->ov_enable = value != 0;
      C->have_topo = false;   // the maps are built by kah_set_topology
      return KAH_OK;
    case KAH_OPT_VEL_TILES:
      if (value != 0 && value != 1) return C->fail(KAH_EINVAL, "velocity tiles must be 0 or 1");
      C->vt_enable = value != 0;
      C->have_topo = false;   // the tile table is built by kah_set_topology
      return KAH_OK;
    case KAH_OPT_BRANCHES:
      if (value < 1 || value > kah_ctx::MAX_BRANCHES)
        return C->fail(KAH_EINVAL, "branches must be 1.." + std::to_string(kah_ctx::MAX_BRANCHES));
      C->branches = int(value);
      C->drop_graphs();
      return KAH_OK;
    case KAH_OPT_REGULAR:
      if (value != 0 && value != 1) return C->fail(KAH_EINVAL, "regular must be 0 or 1");
      C->reg_mode = int(value);
      C->have_topo = false;   // the launch lists are built by kah_set_topology
      return KAH_OK;
    case KAH_OPT_SLOPE_VPT:
      if (value < 1 || value > SLOPE_MAX_VPT)
        return C->fail(KAH_EINVAL, "slope velocity cells per thread must be 1.." +
                                       std::to_string(SLOPE_MAX_VPT));
      C->slope_vpt = int(value);
      C->have_topo = false;   // chunking changes: set the topology again
      return KAH_OK;
    default:
      return C->fail(KAH_EINVAL, "unknown option " + std::to_string(option));
  }
}

int kah_kernel_times(kah_ctx* C, int reset, double* out_ms, int64_t* out_calls) {
  if (!C) return KAH_EINVAL;
  KAH_TRY(C->hip(hipStreamSynchronize(C->stream), "kernel times"));
  for (auto& e : C->events) {
    float ms = 0.f;
    KAH_TRY(C->hip(hipEventElapsedTime(&ms, e.a, e.b), "hipEventElapsedTime"));
    C->t_sum[e.cls] += ms;
    C->t_cnt[e.cls] += 1;
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  C->events.clear();
  for (int k = 0; k < 2; ++k) {
    if (out_ms) out_ms[k] = C->t_cnt[k] ? C->t_sum[k] / double(C->t_cnt[k]) : 0.0;
    if (out_calls) out_calls[k] = C->t_cnt[k];
  }
  if (reset) {
    C->t_sum[0] = C->t_sum[1] = 0.0;
    C->t_cnt[0] = C->t_cnt[1] = 0;
    C->timing = true;
  }
  return KAH_OK;
}

}  // extern "C"

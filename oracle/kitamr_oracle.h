/*
 * kitamr_oracle.h -- CPU FP64 restatement of KitAMR.jl's per-step phase-space update.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the HIP path is compared with;
 * it is linked by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg
 * and by nothing else. The product path never calls it.
 *
 * It consumes the same flattened topology as the device library (include/kitamr_hip.h)
 * and runs the reference's algorithm in the reference's order: per-cell slope dispatch
 * with the two-pointer velocity walks, the per-face scatter of the face list, and the
 * per-cell relaxation update.
 */
#ifndef KITAMR_ORACLE_H
#define KITAMR_ORACLE_H
#include <stdint.h>
#include "../include/kitamr_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

orc_ctx* orc_create(const kah_params* params, const kah_topology* topo);
void orc_destroy(orc_ctx* o);
const char* orc_last_error(const orc_ctx* o);
int orc_set_state(orc_ctx* o, const double* w, const double* prim, const double* df);
int orc_get_state(orc_ctx* o, double* w, double* prim, double* qf, double* sw, double* df,
                  double* sdf);
int orc_get_flux(orc_ctx* o, double* flux, double* vs_flux);
int orc_get_ib(orc_ctx* o, double* aux_df);   /* SolidNeighbor df [face][k][i], last flux */
int orc_set_threads(int n);
int orc_slope(orc_ctx* o);
int orc_flux(orc_ctx* o, double dt);
int orc_iterate(orc_ctx* o, double dt, int64_t step, double* sum_res, double* sum_avg);

/* Point kernels exposed for closed-form tests (lib/KitCore restated). */
void orc_discrete_maxwell(int dim, int ndf, int64_t n, const double* vs_mid,
                          const double* prim, double K, double* out);
void orc_micro_to_macro(int dim, int ndf, int64_t n, const double* vs_mid,
                        const double* df, const double* weight, double* w_out);
void orc_heat_flux(int dim, int ndf, int64_t n, const double* vs_mid, const double* df,
                   const double* weight, const double* prim, double* q_out);
void orc_shakhov_part(int dim, int ndf, int64_t n, const double* vs_mid, const double* F,
                      const double* prim, const double* qf, double Pr, double K,
                      double* out);
/* conserved_I_porjection! of one grid: df ([ndf, n], h row first) is modified in place so
 * that its moments (with b's internal energy, ndf == 2) equal w. */
void orc_i_projection(int dim, int ndf, int64_t n, const double* vs_mid, const double* weight,
                      double* df, const double* w);
/* vs_resolution and the per-velocity-cell refine / coarsen flags of vs_refine! /
 * vs_coarsen! (src/Velocity_space/AMR.jl:24-167, Criteria.jl, Neighbor.jl) on the current
 * state; flags as kah_vs_flags. */
int orc_vs_resolution(orc_ctx* o, double* vr);
int orc_set_sdf(orc_ctx* o, const double* sdf);   /* VsData.sdf of all cells (tests) */
int orc_vs_flags(orc_ctx* o, const kah_vs_adapt* p, const double* vr, uint8_t* flags);
void orc_get_prim(int dim, const double* w, double gamma, double* prim);
void orc_get_conserved(int dim, const double* prim, double gamma, double* w);
double orc_get_tau(int dim, const double* prim, double mu, double omega);
double orc_ref_vhs_vis(double Kn, double omega, double alpha, double T_ref);

#ifdef __cplusplus
}
#endif
#endif

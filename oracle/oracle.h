/*
 * oracle.h — FP64 CPU restatement of the WeldFormGPU per-step SPH pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP product is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (libwfhip) never links or calls
 * it.
 *
 * It restates, in source order and FP64 without FMA contraction (the
 * reference builds with --fmad=false --prec-div=true --prec-sqrt=true,
 * CMakeLists.txt:29), the arithmetic of:
 *   src/Domain.cpp:194-398, 473-574, 989-1005   lattice generators, zones
 *   src/cuNSearch/src/cuNSearchKernels.cu:125-233 FP32 neighbour criterion
 *   src/cuda/Functions.cu:27-283                  cubic-spline kernel, EOS
 *   src/cuda/InteractionAlt.cu:5-383              the three pair loops
 *   src/cuda/tensor3.cu:62-312                    tensor helpers
 *   src/cuda/Mechanical.cu:142-203,289-298,458-589 per-particle kernels
 *   src/cuda/Material.cuh:192-242                 Hollomon / Johnson-Cook
 *   src/cuda/Thermal.cu:16-142                    conduction, convection, leapfrog
 *   src/cuda/Energy.cu:5-19                       internal energy
 *   src/cuda/Geometry.cu:26-88                    free-surface normals and flag
 *   src/cuda/Contact.cu:29-62,85-112,134-321      contact candidates and forces
 *   src/cuda/Mesh.cu:48-276                       plate mesh, rigid motion, normals
 *   src/cuda/SolverFraser.cu:30-409               step order and rebuild policy
 * All arrays are in the caller's (original) particle order.
 */
#ifndef WF_ORACLE_H
#define WF_ORACLE_H

#include <stdint.h>
#include "../include/wf_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

struct or_contact;

typedef struct or_state {
  int n, solid;
  double *x, *v, *a, *u, *va, *xs;                 /* 3n; xs = snapshot x_h */
  double *rho, *rho0, *drho, *m, *h, *p, *Cs, *G;  /* n */
  double *sigma_y, *pl_strain, *sigma_eq, *eff_strain_rate, *int_energy;
  double *T, *Ta, *Tb, *dTdt, *k_T, *cp_T, *q_plastic, *q_friction;
  double *sigma, *shearstress, *strain, *strrate, *rotrate; /* 6n */
  int *ID, *BC_T;
  int *nb_count, *nb_offset; /* n, n+1 */
  int *nb_list;
  int64_t nb_cap;
  /* Newton statistics of the last or_stress_strain call */
  int64_t newton_iters, newton_calls;
  struct or_contact* contact; /* null: no contact data */
  double* xb;                 /* 3n: positions at the last build
                                 (WF_REBUILD_PLASTIC_ZONE; may be null)     */
} or_state;

/* rigid triangle meshes, all meshes concatenated (TriMesh_d, Mesh.cuh:19-54) */
typedef struct or_mesh {
  int nn, ne, n_meshes;
  double *x, *v;          /* 3*nn interleaved: node positions, velocities  */
  int* node_mesh;         /* nn: owning mesh                                */
  int* el;                /* 3*ne global node indices                       */
  double *c, *nrm, *pp;   /* 3*ne centroids, 3*ne unit normals, ne planes   */
  double* motion;         /* 9 per mesh: velocity, angular velocity, centre */
} or_mesh;

typedef struct or_contact {
  or_mesh* mesh;
  double *normal, *cf;    /* 3n: surface normal, contact force              */
  int* idlock;            /* n: ID != 0 at the first surface pass           */
  int *cnb_count, *cnb_list; /* n, 20n: candidates, ascending element index */
  int surf_init;
  int cont_over_cap;
  double totmass;
} or_contact;

typedef struct or_run {
  double time;
  int64_t step;
  int ts_i;
  int is_yielding;
  int isfirst_step;
  int ns_builds;
  double max_pl_strain;
  int nb_over_cap;
  int rebuild_pending;  /* WF_REBUILD_PLASTIC_ZONE: rebuild at the next step */
} or_run;

/* geometry (Domain.cpp) ----------------------------------------------------*/
int or_gen_box(const double V[3], double Lx, double Ly, double Lz, double r,
               double* xyz, int cap, double density, double* mass_out);
int or_gen_cylinder(const double V[3], double Rxy, double Lz, double r,
                    double* xyz, int cap, double density, double* mass_out);
int or_half_part_count(double r, double R, int xinc);
int or_gen_cylinder_polar(int tag, double Rxy, double Lz, double r, double density,
                          double ang, int rows, double r_i, double* xyz, int* ID,
                          int cap, double* mass_out);
int or_assign_zone(const double* xyz, int n, const double start[3],
                   const double end[3], int id, int* ID);

/* neighbour search (FP32 criterion); rows sorted ascending -----------------*/
int64_t or_neighbors(const double* xyz, int n, double radius, int* counts,
                     int* offsets, int* list, int64_t cap, int* over_cap);
int64_t or_neighbors_bruteforce(const double* xyz, int n, double radius,
                                int* counts, int* offsets, int* list,
                                int64_t cap);

/* math helpers ------------------------------------------------------------*/
double or_kernel(double q, double h);
double or_grad_kernel(double q, double h);
double or_jc_yield(const wf_material_params* mp, double ep, double eps_rate,
                   double T);
double or_jc_hardening(const wf_material_params* mp, double ep,
                       double eps_rate, double T);
double or_hollomon_yield(const wf_material_params* mp, double ep);
/* Newton radial return on s_trial - 3G dg - s_y(ep+dg) = 0 for JC. */
double or_jc_newton(const wf_material_params* mp, double s_trial, double G,
                    double ep, double eps_rate, double T, double g0,
                    double tol, int max_iter, int* iters);

/* per-step kernels --------------------------------------------------------*/
void or_calc_dens_inc(or_state* s, const wf_config* c);
void or_update_density(or_state* s, double dt);
void or_calc_rate_tensors(or_state* s, const wf_config* c);
void or_pressure(or_state* s);
void or_stress_strain(or_state* s, const wf_config* c);
void or_calc_accel(or_state* s, const wf_config* c);
void or_apply_axisymm_bc(or_state* s);
void or_apply_bc_vel(or_state* s, int zone, double vx, double vy, double vz);
void or_update_pos_fraser(or_state* s, double dt);
void or_update_vel(or_state* s, double dt);
void or_calc_temp_inc(or_state* s, const wf_config* c);
void or_thermal_calcs(or_state* s, const wf_config* c, int* isfirst_step);
void or_int_energy(or_state* s, const wf_config* c);
double or_int_energy_sum(const or_state* s);

/* neighbour rebuild from the snapshot (or x) into s->nb_* */
int or_nsearch_build(or_state* s, const wf_config* c, int* over_cap);
/* WF_REBUILD_PLASTIC_ZONE: 1 when a particle with eps_p > ep_hot has moved
 * more than ns_skin / 2 from its position at the last build */
int or_plastic_zone_moved(const or_state* s, const wf_config* c);

/* Return-map surrogate network (no reference counterpart): FP64 forward and
 * mean-squared-error gradient of the 4 -> 32 -> 32 -> 1 ReLU MLP, parameter
 * order W1[4][32], b1[32], W2[32][32], b2[32], W3[32], b3 (1249 values). */
double or_mlp_forward(const double* w, const double* x);
double or_mlp_loss_grad(const double* w, const double* X, const double* Y, int n, double* grad);

/* contact ------------------------------------------------------------------*/
/* AxisPlaneMesh(2, false, p1, p2, dens), Mesh.cu:48-194; returns elements */
int or_gen_plane_mesh(const double p1[3], const double p2[3], int dens, double* nodes,
                      int* elnode);
void or_mesh_geom(or_mesh* m);           /* CalcCentroids/CalcNormals/UpdatePlaneCoeff */
void or_mesh_move(or_mesh* m, double dt); /* TriMesh_d::Move                         */
void or_calc_surface(or_state* s, const wf_config* c);  /* CalculateSurface          */
void or_contact_nb(or_state* s, const wf_config* c);    /* CalcContactNb (brute force) */
void or_contact_forces(or_state* s, const wf_config* c);/* CalcContactForcesWang     */

/* OpenMP threads of the oracle's particle loops (the CPU baseline is timed
 * on one thread and on all of them); or_get_threads returns the current cap */
void or_set_threads(int n);
int or_get_threads(void);

/* one Fraser step (SolverFraser.cu:164-399) */
int or_step(or_state* s, const wf_config* c, or_run* run, int nbc,
            const int* bc_zone, const double* bc_val);

#ifdef __cplusplus
}
#endif
#endif

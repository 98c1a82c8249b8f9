/*
 * oracle.h — CPU restatement of aperi-mech's explicit internal-force step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in aperi-mech_amd/.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product never links it.
 *
 * Conventions follow the reference's field layout (include/Field.h:129-138):
 *   node vectors   n_nodes x 3, row-major (x,y,z per node)
 *   tensors        n_elem x 9, row-major (component 3*i+j)
 *   connectivity   n_elem x nodes_per_elem
 *
 * Pinning status of each function (what fixes it independent of this code)
 * is listed in the header comment of the .c file that defines it.  Functions
 * whose values are pinned only at rounding level (never bit level):
 *   - Eigen 3x3 inverse / FullPivLU (third party; tolerance-level parity)
 *   - summation order of the reference's atomic scatters
 */
#ifndef APERI_ORACLE_H
#define APERI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_TET4 = 0, OR_HEX8 = 1 };
enum { OR_LINEAR_ELASTIC = 0, OR_NEO_HOOKEAN = 1, OR_J2 = 2, OR_SVK = 3, OR_NH_DAMAGE = 4, OR_CREEP = 5, OR_DP = 6 };

/* Material: uniform parameters, optionally overridden per element. */
typedef struct {
    int32_t kind;
    int32_t state_per_qp;      /* Hex8 + J2 only: 1 = one state slot per Gauss point (extension) */
    double lambda, mu;         /* Lame parameters */
    double yield_stress, hardening_modulus;
    const double *elem_lambda; /* optional [n_elem] */
    const double *elem_mu;     /* optional [n_elem] */
    /* neo-Hookean with damage (include/Materials/NeoHookeWithDamage.h) */
    double I1_critical, I1_failure, alpha, maximum_damage, maximum_damage_increment;
    /* power-law creep (include/Materials/PowerLawCreep.h): bulk modulus, A, n, m and the
     * constant temperature */
    double bulk, creep_A, creep_n, creep_m, temperature;
    /* Drucker-Prager (include/Materials/DruckerPrager.h; bulk as above) */
    double dp_A1, dp_A2, dp_A2G;
    /* rate-form materials: H and P of the previous step ([n_elem][9], row-major; owned
     * by the caller, or_solve refreshes them each step) and the increment */
    double *H_n, *P_n;
    double dt;
    /* updated Lagrangian (1; semi-Lagrangian 2): the coordinates passed are those at the
     * start of the step (semi: the reference configuration), the displacements the step's
     * increments (semi: accumulated since the reference update), H_n the previous (semi:
     * the reference) displacement gradient */
    int32_t updated;
    int32_t pad_;
} or_material;

/* 1 for materials that read the previous step's H and P (velocity gradient, old stress). */
int or_rate_form(int kind);
/* The rate-form stress of `mat` from H, the previous step's Hn, Pn and the increment dt. */
void or_stress_rate(const or_material *mat, double lambda, double mu, const double H[9], const double Hn[9],
                    const double Pn[9], double dt, const double *state_n, double *state_np1, double P[9]);

/* State values per evaluation point: J2 1 (equivalent plastic strain), neo-Hookean with
 * damage 2 (damage, I1), others 0. */
int or_state_width(int kind);

/* ---- math (or_math.c) ---- */
int or_lame_from_E_nu(double E, double nu, double *lambda, double *mu, double *K);
double or_det_ap_im1(const double A[9]);
double or_kernel_value(const double d[3], double R);
void or_inverse3_cofactor(const double A[9], double Ainv[9], double *det);
int or_fullpivlu_inverse(int n, const double *A, double *Ainv);
int or_elastic_properties(int given_a, double a, int given_b, double b, double out[5]);

/* ---- materials (or_material.c) ---- */
/* Stress of the material `mat` with Lame parameters lambda, mu (per element);
 * state_n / state_np1 hold or_state_width(kind) values (NULL when 0). */
void or_stress_mat(const or_material *mat, double lambda, double mu, const double H[9], const double *state_n,
                   double *state_np1, double P[9]);
void or_stress(int kind, double lambda, double mu, double yield_stress, double hardening,
               const double H[9], const double *state_n, double *state_np1, double P[9]);

/* ---- element kinematics (or_element.c) ---- */
void or_tet4_bmatrix(const double X[12], double B[12], double *w);
void or_hex8_bmatrix(const double X[24], int gauss_id, double B[24], double *w);
void or_smoothed_tet4_bmatrix(const double X[12], double B[12], double *V);
void or_smoothed_hex8_bmatrix(const double X[24], double B[24], double *V);
void or_element_volumes(int etype, int64_t n_elem, const int32_t *conn, const double *X, double *vol);

/* Gaussian internal force over all elements.  f is ZEROED then accumulated.
 * state arrays may be NULL for stateless materials; for J2 they hold one value
 * per element (or 8 per element with state_per_qp).  H, P may be NULL. */
int or_internal_force(int etype, int64_t n_nodes, const double *X, int64_t n_elem, const int32_t *conn,
                      const or_material *mat, const double *u, const double *state_n, double *state_np1,
                      double *H, double *P, double *f, int nthreads);

/* ---- mass and time integration (or_solver.c) ---- */
int or_lumped_mass(int etype, int64_t n_nodes, const double *X, int64_t n_elem, const int32_t *conn,
                   double density, const double *elem_density, double *mass);

/* Problem description for the explicit solver loop (src/Solvers/Explicit.cpp:383-452). */
typedef struct {
    int32_t etype;        /* OR_TET4 / OR_HEX8, or -1 for RK smoothed cells */
    int64_t n_nodes;
    const double *X;
    int64_t n_elem;
    const int32_t *conn;
    or_material mat;
    const double *mass;            /* [n_nodes] one scalar per node */
    const uint8_t *fixed;          /* [n_nodes*3] 1 = velocity/accel held at 0, or NULL */
    const double *node_dt;         /* optional [n_nodes] per-node time increment (ensembles) */
    /* RK smoothed cells (etype == -1) */
    const int64_t *cell_start;     /* [n_elem+1] */
    const int32_t *cell_nodes;     /* [nnz] */
    const double *cell_dphi;       /* [nnz*3] */
    const double *cell_volume;     /* [n_elem] */
    /* two-pass RK force (etype == -1, rk_two_pass = 1): the cells list the element's
     * own nodes; u_phys = Phi d before the cell force, f = Phi^T f_local after it */
    int32_t rk_two_pass;
    int32_t max_neighbors;
    const int32_t *num_neighbors;  /* [n_nodes] */
    const int32_t *neighbors;      /* [n_nodes*max_neighbors] */
    const double *phi;             /* [n_nodes*max_neighbors] */
    /* updated Lagrangian (mat.updated = 1): current coordinates at the start of the step and
     * the step's displacement increments, [n_nodes*3], owned by the caller */
    double *x_cur;
    double *d_inc;
    /* semi-Lagrangian: reference update interval and the caller's step counter */
    int64_t semi_interval;
    int64_t *semi_steps;
    /* F-bar over nodal cells (etype == -1, fbar = 1): the cells above are the subcells,
     * grouped into n_fbar_cells cells of consecutive subcells (fbar_cell_start[n+1]) */
    int32_t fbar;
    int32_t pad2_;
    int64_t n_fbar_cells;
    const int64_t *fbar_cell_start;
    const double *fbar_cell_volume;
    /* nodes that carry degrees of freedom (NULL: all); the others are never updated */
    const uint8_t *active;
} or_problem;

/* Initial force and acceleration (Explicit.cpp:363-376), then n_steps of the
 * central-difference loop.  u, v are inout; a, f, H, P outputs; state_a/state_b
 * are the N/NP1 state buffers (swapped each step); on return *state_out_is_b
 * tells which buffer holds the final NP1 state.  dt is the fixed increment
 * (DirectTimeStepper); with node_dt set, node_dt overrides it per node. */
int or_solve(const or_problem *p, double dt, int n_steps, int do_initial,
             double *u, double *v, double *a, double *f, double *H, double *P,
             double *state_a, double *state_b, int *state_out_is_b, int nthreads);

/* Power-method stable time increment (include/PowerMethodProcessor.h:331-415), fixed
 * iteration count, perturbation eps v about u = 0. */
double or_power_method_dt(const or_problem *p, int n_iter, double epsilon, uint64_t seed, int nthreads);
/* The reference's power method (include/PowerMethodProcessor.h:331-470) about u_in with
 * per-node max edge lengths, its convergence tests and its result rule; previous_dt seeds
 * the first convergence test, `eigenvector` (3 n_nodes, optional) carries the vector from
 * one call to the next (used as the start when eigenvector_valid); returns 0 on success. */
int or_power_method(const or_problem *p, const double *u_in, const double *state_n, const double *max_edge,
                    int n_iter, double epsilon, double tolerance, double variation_tolerance, uint64_t seed,
                    double previous_dt, double *eigenvector, int eigenvector_valid, int nthreads, double *dt_out,
                    int *iterations_out, int *converged_out);

/* ---- RKPM (or_rk.c) ---- */
int or_max_edge_length(int64_t n_nodes, const double *X, int64_t n_elem, const int32_t *conn, int npe,
                       const uint8_t *active, double *max_edge);
int or_neighbor_search(int64_t n_nodes, const double *X, const double *radius, int max_neighbors,
                       int brute_force, int32_t *num_neighbors, int32_t *neighbors, int64_t *n_truncated);
int or_rk_function_values(int64_t n_nodes, const double *X, const double *radius, int max_neighbors,
                          const int32_t *num_neighbors, const int32_t *neighbors, double *phi);
/* npe = 4 (Tet4 background) or 8 (Hex8 background): one cell per element */
int64_t or_rk_cells_count(int64_t n_elem, const int32_t *conn, int npe, int max_neighbors,
                          const int32_t *num_neighbors, const int32_t *neighbors, int64_t *cell_start);
int or_rk_cells_build(int64_t n_nodes, const double *X, int64_t n_elem, const int32_t *conn, int npe, int max_neighbors,
                      const int32_t *num_neighbors, const int32_t *neighbors, const double *phi,
                      const int64_t *cell_start, int32_t *cell_nodes, double *cell_dphi, double *cell_volume);
int or_cell_internal_force(int64_t n_nodes, int64_t n_cells, const int64_t *cell_start, const int32_t *cell_nodes,
                           const double *cell_dphi, const double *cell_volume, const or_material *mat,
                           const double *u, const double *state_n, double *state_np1,
                           double *H, double *P, double *f, int nthreads);
int or_rk_mass(int64_t n_nodes, int64_t n_elem, const int32_t *conn, int npe, const double *elem_volume, double density,
               int max_neighbors, const int32_t *num_neighbors, const int32_t *neighbors, const double *phi,
               double *mass);
/* Two-pass method.  Cells of the element's own nodes: start[e] = npe e, grad N_k = (B_k V) / V. */
int or_rk_cells_two_pass(int64_t n_elem, const double *X, const int32_t *conn, int npe, int64_t *cell_start,
                         int32_t *cell_nodes, double *cell_dphi, double *cell_volume);
/* dst_i = sum_k phi_ik src_(nbr_ik), k descending; src_i for a node without neighbors (3 components). */
int or_rk_gather(int64_t n_nodes, int max_neighbors, const int32_t *num_neighbors, const int32_t *neighbors,
                 const double *phi, const double *src, double *dst);
/* dst (zeroed here) gets phi_ik src_i added at every neighbor nbr_ik, ascending i then k; a node
 * without neighbors takes src_i alone (3 components). */
int or_rk_scatter(int64_t n_nodes, int max_neighbors, const int32_t *num_neighbors, const int32_t *neighbors,
                  const double *phi, const double *src, double *dst);

/* ---- nodal strain smoothing on thex meshes, subcells, F-bar (or_nodal.c) ---- */
/* active[i] = 1 for the node of lowest id (node_id[i], or i when NULL) of some hex; returns
 * the number of hexes whose count of active nodes is not exactly one (0 = a valid thex mesh). */
int64_t or_nodal_active(int64_t n_nodes, int64_t n_hex, const int32_t *hexes, const int64_t *node_id,
                        uint8_t *active);
/* hex_cell[e] = rank of the hex's active node among the active nodes (ascending index);
 * returns the number of cells. */
int64_t or_nodal_cells(int64_t n_nodes, int64_t n_hex, const int32_t *hexes, const uint8_t *active,
                       int32_t *hex_cell);
/* ball search whose candidates are only the nodes with range_mask[j] != 0 (NULL: all) */
int or_neighbor_search_masked(int64_t n_nodes, const double *X, const double *radius, int max_neighbors,
                              const uint8_t *range_mask, int brute_force, int32_t *num_neighbors,
                              int32_t *neighbors, int64_t *n_truncated);
/* subcells of several elements: sub_start[n_sub+1] ranges into sub_elems */
int64_t or_rk_subcells_count(int64_t n_sub, const int64_t *sub_start, const int32_t *sub_elems, const int32_t *conn,
                             int npe, int max_neighbors, const int32_t *num, const int32_t *nbr, int64_t *start);
int or_rk_subcells_build(int64_t n_nodes, const double *X, int64_t n_sub, const int64_t *sub_start,
                         const int32_t *sub_elems, const int32_t *conn, int npe, int max_neighbors,
                         const int32_t *num, const int32_t *nbr, const double *phi, const int64_t *start,
                         int32_t *nodes, double *dphi, double *sub_volume, double *elem_volume);
/* F-bar force of cells made of consecutive subcells (cell_sub[n_cells+1]); H, Hbar, P, Pbar
 * per subcell (row-major, any may be NULL); the state per subcell. */
int or_fbar_internal_force(int64_t n_nodes, int64_t n_cells, const int64_t *cell_sub, const double *cell_volume,
                           const int64_t *sub_start, const int32_t *sub_nodes, const double *sub_dphi,
                           const double *sub_volume, const or_material *mat, const double *u, const double *state_n,
                           double *state_np1, double *H, double *Hbar, double *P, double *Pbar, double *f,
                           int nthreads);

#ifdef __cplusplus
}
#endif
#endif

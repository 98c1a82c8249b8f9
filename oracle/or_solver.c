/*
 * or_solver.c — lumped mass and the explicit central-difference loop.
 * TEST INFRASTRUCTURE (oracle).
 *
 * Restates:
 *   lumped mass        src/MassUtils.cpp:26-39 (m_I += rho V_e / n_nodes), :79-112
 *   solver loop        src/Solvers/Explicit.cpp:308-452 (initial force + a, then per step:
 *                      v1/2 = v + dt/2 a; velocity BC; u += dt v1/2; f(u); a = f/m;
 *                      acceleration BC; v = v1/2 + dt/2 a)
 *   node functors      include/ExplicitTimeIntegrator.h:116-126, :267-276, :308-318, :348-357
 *   N/NP1 state swap   Explicit.cpp:393 (UpdateFieldStates) for the element state field
 *   active nodes       the integrator runs over the active part only (Explicit.cpp:316,
 *                      ExplicitTimeIntegrator.h:376-379): nodes without DOFs keep their values
 *   updated Lagrangian ExplicitTimeIntegratorUpdated (include/ExplicitTimeIntegrator.h:193-226,
 *                      :466-486): increment dt v1/2, d += increment, x_{n+1} = x_n + increment,
 *                      the force over x_n with the increments and H_n; semi-Lagrangian
 *                      ExplicitTimeIntegratorSemi (:488-546): increments accumulate, the force
 *                      over the reference configuration with the reference gradient, which
 *                      StepScheduler(interval, interval) refreshes after steps n = k interval
 *                      (src/Application.cpp:301-307, src/Solvers/Explicit.cpp:195-212)
 *
 * Pinning (tests/test_oracle_solver.py): sum m = rho V (MassUtilsTest.cpp:202-265);
 * free flight (zero stiffness path: rigid translation gives f = 0 and u = v0 t exactly);
 * a one-element spring-mass vibration against the closed-form central-difference
 * recurrence; linear-elastic energy conservation to O(dt^2) over 100 steps.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

int or_lumped_mass(int etype, int64_t n_nodes, const double *X, int64_t n_elem, const int32_t *conn,
                   double density, const double *elem_density, double *mass) {
    const int npe = etype == OR_TET4 ? 4 : 8;
    double *vol = (double *)malloc(sizeof(double) * (size_t)(n_elem > 0 ? n_elem : 1));
    if (!vol) return -1;
    or_element_volumes(etype, n_elem, conn, X, vol);
    memset(mass, 0, sizeof(double) * n_nodes);
    for (int64_t e = 0; e < n_elem; ++e) {
        const double rho = elem_density ? elem_density[e] : density;
        const double m = vol[e] * rho;
        const double mn = m / npe;
        for (int k = 0; k < npe; ++k) mass[conn[e * npe + k]] += mn;
    }
    free(vol);
    return 0;
}

static int compute_force(const or_problem *p, const double *u, const double *sn, double *snp1, double *H,
                         double *P, double *f, int nthreads) {
    if (p->etype == -1 && p->rk_two_pass) {  /* ExplicitSolver::ComputeForce (Explicit.cpp:148-186) */
        double *up = (double *)malloc(sizeof(double) * 3 * (size_t)(p->n_nodes > 0 ? p->n_nodes : 1));
        double *fl = (double *)malloc(sizeof(double) * 3 * (size_t)(p->n_nodes > 0 ? p->n_nodes : 1));
        int rc = (up && fl) ? 0 : -1;
        if (!rc) rc = or_rk_gather(p->n_nodes, p->max_neighbors, p->num_neighbors, p->neighbors, p->phi, u, up);
        if (!rc)
            rc = or_cell_internal_force(p->n_nodes, p->n_elem, p->cell_start, p->cell_nodes, p->cell_dphi,
                                        p->cell_volume, &p->mat, up, sn, snp1, H, P, fl, nthreads);
        if (!rc) rc = or_rk_scatter(p->n_nodes, p->max_neighbors, p->num_neighbors, p->neighbors, p->phi, fl, f);
        free(up);
        free(fl);
        return rc;
    }
    if (p->etype == -1 && p->fbar) /* nodal cells with F-bar (ComputeInternalForceSmoothedCell.h:120-413) */
        return or_fbar_internal_force(p->n_nodes, p->n_fbar_cells, p->fbar_cell_start, p->fbar_cell_volume,
                                      p->cell_start, p->cell_nodes, p->cell_dphi, p->cell_volume, &p->mat, u, sn,
                                      snp1, H, NULL, P, NULL, f, nthreads);
    if (p->etype == -1)
        return or_cell_internal_force(p->n_nodes, p->n_elem, p->cell_start, p->cell_nodes, p->cell_dphi,
                                      p->cell_volume, &p->mat, u, sn, snp1, H, P, f, nthreads);
    return or_internal_force(p->etype, p->n_nodes, p->X, p->n_elem, p->conn, &p->mat, u, sn, snp1, H, P, f,
                             nthreads);
}

int or_solve(const or_problem *p, double dt, int n_steps, int do_initial, double *u, double *v, double *a,
             double *f, double *H, double *P, double *state_a, double *state_b, int *state_out_is_b,
             int nthreads) {
    const int64_t N = p->n_nodes;
    const int nt = nthreads > 0 ? nthreads : 1;
    double *sn = state_a, *snp1 = state_b;
    int rc;
    /* rate-form materials: this call's increment; H_n, P_n are the caller's buffers */
    or_problem q = *p;
    q.mat.dt = dt;
    const int rate = or_rate_form(p->mat.kind), ul = p->mat.updated, semi = ul == 2;
    if ((rate || ul) && (!q.mat.H_n || !q.mat.P_n)) return -1;
    if (ul && (!q.x_cur || !q.d_inc)) return -1;
    if (semi && (rate || !q.semi_steps || q.semi_interval < 1)) return -1;
    if (semi && do_initial) *q.semi_steps = 0;
    double *x_np1 = NULL;
    if (ul) {
        x_np1 = (double *)malloc(sizeof(double) * 3 * (size_t)(N > 0 ? N : 1));
        if (!x_np1) return -1;
        q.X = q.x_cur; /* the force integrates over the start-of-step configuration */
    }
    p = &q;
    if (do_initial) {
        /* Explicit.cpp:363-376: force and acceleration at the initial state (written to NP1) */
        rc = compute_force(p, ul ? q.d_inc : u, snp1, snp1, H, P, f, nt);
        if (rc) { free(x_np1); return rc; }
#pragma omp parallel for schedule(static) num_threads(nt)
        for (int64_t i = 0; i < N; ++i) {
            if (p->active && !p->active[i]) continue;
            for (int d = 0; d < 3; ++d) a[3 * i + d] = f[3 * i + d] / p->mass[i];
        }
    }
    for (int s = 0; s < n_steps; ++s) {
        /* UpdateFieldStates: N <- NP1 */
        double *t = sn; sn = snp1; snp1 = t;
        if (rate || (ul && !semi)) {
            memcpy(q.mat.H_n, H, sizeof(double) * 9 * (size_t)p->n_elem);
            memcpy(q.mat.P_n, P, sizeof(double) * 9 * (size_t)p->n_elem);
        }
#pragma omp parallel for schedule(static) num_threads(nt)
        for (int64_t i = 0; i < N; ++i) {
            if (p->active && !p->active[i]) continue; /* nodes without DOFs are not integrated */
            const double dti = p->node_dt ? p->node_dt[i] : dt;
            const double h = 0.5 * dti;
            for (int d = 0; d < 3; ++d) v[3 * i + d] = v[3 * i + d] + h * a[3 * i + d];
        }
        if (p->fixed)
            for (int64_t i = 0; i < 3 * N; ++i)
                if (p->fixed[i]) v[i] = 0.0;
        if (ul) {
            for (int64_t i = 0; i < 3 * N; ++i) {
                const double inc = dt * v[i];
                u[i] = u[i] + inc;
                if (semi) {
                    q.d_inc[i] = q.d_inc[i] + inc;
                } else {
                    q.d_inc[i] = inc;
                    x_np1[i] = q.x_cur[i] + inc;
                }
            }
        } else {
#pragma omp parallel for schedule(static) num_threads(nt)
            for (int64_t i = 0; i < N; ++i) {
                if (p->active && !p->active[i]) continue;
                const double dti = p->node_dt ? p->node_dt[i] : dt;
                for (int d = 0; d < 3; ++d) u[3 * i + d] = u[3 * i + d] + dti * v[3 * i + d];
            }
        }
        rc = compute_force(p, ul ? q.d_inc : u, sn, snp1, H, P, f, nt);
        if (rc) { free(x_np1); return rc; }
        if (ul && !semi) memcpy(q.x_cur, x_np1, sizeof(double) * 3 * (size_t)N);
#pragma omp parallel for schedule(static) num_threads(nt)
        for (int64_t i = 0; i < N; ++i) {
            if (p->active && !p->active[i]) continue;
            for (int d = 0; d < 3; ++d) a[3 * i + d] = f[3 * i + d] / p->mass[i];
        }
        if (p->fixed)
            for (int64_t i = 0; i < 3 * N; ++i)
                if (p->fixed[i]) a[i] = 0.0;
#pragma omp parallel for schedule(static) num_threads(nt)
        for (int64_t i = 0; i < N; ++i) {
            if (p->active && !p->active[i]) continue;
            const double dti = p->node_dt ? p->node_dt[i] : dt;
            const double h = 0.5 * dti;
            for (int d = 0; d < 3; ++d) v[3 * i + d] += h * a[3 * i + d];
        }
        if (semi && ++*q.semi_steps % q.semi_interval == 0) { /* UpdateReferenceConfiguration */
            for (int64_t i = 0; i < 3 * N; ++i) {
                q.x_cur[i] = q.x_cur[i] + q.d_inc[i];
                q.d_inc[i] = 0.0;
            }
            memcpy(q.mat.H_n, H, sizeof(double) * 9 * (size_t)p->n_elem);
        }
    }
    if (state_out_is_b) *state_out_is_b = (snp1 == state_b);
    free(x_np1);
    return 0;
}

static uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

/*
 * Stable time increment by power iteration on M^-1 K, with K applied as the
 * finite difference of the internal force (include/PowerMethodProcessor.h:
 * 331-415): v <- -(f(u + eps v) - f(u)) / (eps m), lambda = |v|, dt = 2/sqrt(lambda).
 * Runs n_iter iterations from a seeded start vector; fixed DOFs are zeroed.
 */
double or_power_method_dt(const or_problem *p, int n_iter, double epsilon, uint64_t seed, int nthreads) {
    const int64_t N = p->n_nodes, M = 3 * N;
    double *vec = (double *)calloc((size_t)M, sizeof(double));
    double *up = (double *)calloc((size_t)M, sizeof(double));
    double *f0 = (double *)calloc((size_t)M, sizeof(double));
    double *f1 = (double *)calloc((size_t)M, sizeof(double));
    const int64_t ns = or_state_width(p->mat.kind) > 0
                           ? or_state_width(p->mat.kind) * (p->etype == OR_HEX8 && p->mat.state_per_qp ? 8 : 1) * p->n_elem
                           : 1;
    double *st = (double *)calloc((size_t)ns, sizeof(double));
    double *st2 = (double *)calloc((size_t)ns, sizeof(double));
    double lambda = 0.0;
    if (!vec || !up || !f0 || !f1 || !st || !st2) goto done;
    for (int64_t i = 0; i < M; ++i) {
        vec[i] = (double)(splitmix64(seed ^ (uint64_t)i) >> 11) * (1.0 / 9007199254740992.0) - 0.5;
        if (p->fixed && p->fixed[i]) vec[i] = 0.0;
    }
    {
        double nrm = 0.0;
        for (int64_t i = 0; i < M; ++i) nrm += vec[i] * vec[i];
        nrm = sqrt(nrm);
        for (int64_t i = 0; i < M; ++i) vec[i] /= nrm;
    }
    compute_force(p, up, st, st2, NULL, NULL, f0, nthreads);
    for (int it = 0; it < n_iter; ++it) {
        for (int64_t i = 0; i < M; ++i) up[i] = epsilon * vec[i];
        compute_force(p, up, st, st2, NULL, NULL, f1, nthreads);
        double nrm = 0.0;
        for (int64_t i = 0; i < M; ++i) {
            double w = -(f1[i] - f0[i]) / (epsilon * p->mass[i / 3]);
            if (p->fixed && p->fixed[i]) w = 0.0;
            vec[i] = w;
            nrm += w * w;
        }
        nrm = sqrt(nrm);
        lambda = nrm;
        for (int64_t i = 0; i < M; ++i) vec[i] /= nrm;
    }
done:
    free(vec); free(up); free(f0); free(f1); free(st); free(st2);
    return lambda > 0.0 ? 2.0 / sqrt(lambda) : 0.0;
}

/*
 * The reference's power-method stable increment, restated step by step
 * (include/PowerMethodProcessor.h:331-470 ComputeStableTimeIncrement,
 *  InitializeEigenvector :195-221, PerturbDisplacementCoefficients :230-260,
 *  ComputeNextEigenvector :262-330; NormalizeField = L2 norm,
 *  include/EntityProcessor.h:676-697):
 *   v_i = U[-1,1] * eps * l_i                      (start; l = max edge length per node)
 *   repeat: u_eps = u + v l eps (fixed DOFs: u); F1 = F(u_eps)
 *           v = (F1 - F(u)) / (l eps m) (fixed DOFs: 0); lambda = |v|_2; v /= lambda
 *           dt = 2 / sqrt(lambda); stop when |dt - dt_prev| < tol dt, or when the variance
 *           of the last five dt is below vtol x their mean
 *   result: the last dt if converged, else the smallest of the last five.
 * The first comparison is against previous_dt (the caller's last full increment, or
 * 1e-16 when it is 0: :363-364).  The eigenvector persists between calls of one
 * processor (:350-361, :468-469): with eigenvector_valid the iteration starts from
 * `eigenvector` (the normalised vector the previous call ended with), and the final
 * vector is written back to it when it is given.
 * The random start is splitmix64(seed ^ dof) (the reference draws Kokkos XorShift64
 * numbers keyed by node id: the start differs, the converged eigenvalue does not;
 * parity of the start vector itself is unpinned).  The reference's oscillation test
 * (:424-437) takes its else branch at i = 0 and never reports oscillation; restated.
 */
int or_power_method(const or_problem *p, const double *u_in, const double *state_n, const double *max_edge,
                    int n_iter, double epsilon, double tolerance, double variation_tolerance, uint64_t seed,
                    double previous_dt, double *eigenvector, int eigenvector_valid, int nthreads, double *dt_out,
                    int *iterations_out, int *converged_out) {
    const int64_t N = p->n_nodes, M = 3 * N;
    const int64_t ns = or_state_width(p->mat.kind) > 0
                           ? or_state_width(p->mat.kind) * (p->etype == OR_HEX8 && p->mat.state_per_qp ? 8 : 1) * p->n_elem
                           : 1;
    double *vec = (double *)calloc((size_t)M, sizeof(double));
    double *up = (double *)calloc((size_t)M, sizeof(double));
    double *f0 = (double *)calloc((size_t)M, sizeof(double));
    double *f1 = (double *)calloc((size_t)M, sizeof(double));
    double *sn = (double *)calloc((size_t)ns, sizeof(double));
    double *snp1 = (double *)calloc((size_t)ns, sizeof(double));
    int rc = -1;
    if (!vec || !up || !f0 || !f1 || !sn || !snp1) goto done;
    if (state_n && or_state_width(p->mat.kind) > 0) memcpy(sn, state_n, sizeof(double) * (size_t)ns);
    if (eigenvector && eigenvector_valid) {
        memcpy(vec, eigenvector, sizeof(double) * (size_t)M);
    } else {
        for (int64_t i = 0; i < M; ++i) {
            const double r = (double)(splitmix64(seed ^ (uint64_t)i) >> 11) * (1.0 / 9007199254740992.0);
            vec[i] = (2.0 * r - 1.0) * (epsilon * max_edge[i / 3]);
        }
    }
    if ((rc = compute_force(p, u_in, sn, snp1, NULL, NULL, f0, nthreads)) != 0) goto done;
    {
        enum { W = 5 };
        double recent[W] = {0, 0, 0, 0, 0};
        double dt_np1 = previous_dt == 0.0 ? 1.0e-16 : previous_dt, dt_n;
        int converged = 0, oscillating = 0, used = n_iter;
        for (int k = 0; k < n_iter; ++k) {
            dt_n = dt_np1;
            for (int64_t i = 0; i < M; ++i)
                up[i] = (p->fixed && p->fixed[i]) ? u_in[i] : fma(vec[i] * max_edge[i / 3], epsilon, u_in[i]);
            if ((rc = compute_force(p, up, sn, snp1, NULL, NULL, f1, nthreads)) != 0) goto done;
            double s = 0.0;
            for (int64_t i = 0; i < M; ++i) {
                const double w = (p->fixed && p->fixed[i]) ? 0.0
                                                             : (f1[i] - f0[i]) / ((max_edge[i / 3] * epsilon) * p->mass[i / 3]);
                vec[i] = w;
                s = fma(w, w, s);
            }
            const double lambda = sqrt(s);
            for (int64_t i = 0; i < M; ++i) vec[i] /= lambda;
            dt_np1 = 2.0 / sqrt(lambda);
            recent[k % W] = dt_np1;
            if (fabs(dt_np1 - dt_n) < tolerance * dt_np1) {
                converged = 1;
                used = k + 1;
                break;
            }
            if (k >= W) {
                double mean = 0.0, var = 0.0;
                for (int i = 0; i < W; ++i) mean += recent[i];
                mean /= W;
                for (int i = 0; i < W; ++i) var += (recent[i] - mean) * (recent[i] - mean);
                var /= (W - 1);
                if (k > W + 1) {
                    double alt = 0.0;
                    for (int i = 0; i < W - 1; ++i) {
                        const double diff = recent[i + 1] - recent[i];
                        if (i > 0 && diff * alt < 0) {
                            oscillating = 1;
                        } else {
                            oscillating = 0;
                            break;
                        }
                        alt = diff;
                    }
                }
                if (var < variation_tolerance * mean || (oscillating && var < variation_tolerance * mean * 2.0)) {
                    converged = 1;
                    used = k + 1;
                    break;
                }
            }
        }
        double wmin = recent[0];
        for (int i = 1; i < W; ++i) wmin = recent[i] < wmin ? recent[i] : wmin;
        *dt_out = converged ? (oscillating ? wmin : dt_np1) : wmin;
        if (eigenvector) memcpy(eigenvector, vec, sizeof(double) * (size_t)M);
        if (iterations_out) *iterations_out = used;
        if (converged_out) *converged_out = converged;
        rc = 0;
    }
done:
    free(vec); free(up); free(f0); free(f1); free(sn); free(snp1);
    return rc;
}

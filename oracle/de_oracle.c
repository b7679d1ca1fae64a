/*
 * CPU oracle for the DE (discontinuous-elevation) shallow-water step.
 *
 * TEST INFRASTRUCTURE ONLY. This file restates, in plain C, the reference's
 * default (mode 0) FP64 kernels so that the HIP path can be checked against it.
 * Only tests, the smoke check and the CPU-baseline leg of the benchmark load it;
 * the product library never links or calls it.
 *
 * Each function cites the reference lines it follows
 * (paths relative to the anuga_core tree):
 *   oracle_rotate               anuga/shallow_water/sw_domain_orig.c:37-57
 *   oracle_flux_function_central sw_domain_orig.c:284-505
 *   oracle_compute_fluxes       sw_domain_orig.c:765-1150 (edge computed once,
 *                               mirrored to the neighbour; riverwall edges
 *                               counted in loop order as at :828-837, :880-893)
 *   oracle_compute_fluxes_lts   the same loop with local flux updating's skip
 *                               test and static timestep (:784-829)
 *   oracle_compute_flux_update_frequency sw_domain_orig.c:509-669
 *   oracle_compute_fluxes_gather same arithmetic, each triangle evaluating its
 *                               own three edges (OpenMP-parallel, deterministic;
 *                               riverwall index from edge_river_wall_counter as
 *                               in sw_domain_openacc.c:586)
 *   oracle_adjust_edgeflux_with_weir sw_domain_orig.c:671-763 (with the call
 *                               site's conditions, :905-968)
 *   oracle_protect              sw_domain_orig.c:1215-1281
 *   oracle_extrapolate          sw_domain_orig.c:1994-2680 (+ :1286-1333)
 *   oracle_manning_friction_flat sw_domain_orig.c:2805-2838
 *   oracle_manning_friction_sloped sw_domain_orig.c:2840-2888 (+ _gradient,
 *                               anuga/utilities/util_ext.h:52-92)
 *   oracle_update               anuga/abstract_2d_finite_volumes/quantity.c:772-822
 *   oracle_fix_negative_cells   sw_domain_orig.c:2890-2910
 *   oracle_saxpy / oracle_backup quantity.c:735-769
 *
 * Pinning: every function is compared bit for bit against the reference's own
 * compiled C (oracle/_ref) on seeded states, the flux function against the
 * reference's known-answer tests, and the orchestration against the
 * reference's golden trajectories (tests/test_oracle.py).
 *
 * Build: gcc -O2 -fno-fast-math -ffp-contract=off (no FMA contraction: the
 * reference's default build has none).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

/*
 * The reference's two library pow calls: pow(h, 7/3) of Manning friction and
 * pow(1 - Q2/Q1, 0.385) of the weir, made exactly as the reference makes them
 * (the C library's pow; glibc's on this platform).
 */
double oracle_pow(double x, double y) { return pow(x, y); }

double oracle_pow_7_3(double h) { return oracle_pow(h, 7.0 / 3.0); }

typedef struct {
    int64_t N;                       /* triangles */
    int64_t M;                       /* boundary edges */
    const int64_t *neighbours;       /* (N,3), boundary edge -> -(j+1) */
    const int64_t *neighbour_edges;  /* (N,3) */
    const int64_t *surrogate;        /* (N,3) */
    const int64_t *nboundaries;      /* (N) */
    const int64_t *tri_full_flag;    /* (N) */
    const double *normals;           /* (N,6) */
    const double *edgelengths;       /* (N,3) */
    const double *radii;             /* (N) */
    const double *areas;             /* (N) */
    const double *centroids;         /* (N,2) */
    const double *edge_midpoints;    /* (N,6) */
    /* riverwalls (edge_flux_type NULL: none). Flagged half-edges are ranked
     * in (triangle, edge) order; rank r (0-based) owns riverwall_elevation[r]
     * and hydraulic-properties row riverwall_rowIndex[r] (columns Qfactor, s1,
     * s2, h1, h2). edge_river_wall_counter holds rank + 1 (0: not a wall). */
    const int64_t *edge_flux_type;          /* (N,3) 1 = riverwall */
    const int64_t *edge_river_wall_counter; /* (N,3) */
    const double *riverwall_elevation;      /* (R) */
    const int64_t *riverwall_rowIndex;      /* (R) */
    int64_t ncol_riverwall_hydraulic_properties;
    const double *riverwall_hydraulic_properties;
} oracle_mesh;

typedef struct {
    double g;
    double epsilon;
    double minimum_allowed_height;
    double beta_w, beta_w_dry, beta_uh, beta_uh_dry, beta_vh, beta_vh_dry;
    int64_t low_froude;
    int64_t extrapolate_velocity_second_order;
} oracle_params;

typedef struct {
    double *w, *uh, *vh, *z, *h;            /* centroid (N) */
    double *w_e, *uh_e, *vh_e, *z_e, *h_e;  /* edge (3N) */
    double *w_v, *uh_v, *vh_v, *z_v, *h_v;  /* vertex (3N) */
    double *w_b, *uh_b, *vh_b;              /* boundary (M) */
    double *ex_w, *ex_uh, *ex_vh;           /* explicit update (N) */
    double *si_w, *si_uh, *si_vh;           /* semi-implicit update (N) */
    double *xwork, *ywork;                  /* (N) */
    double *max_speed;                      /* (N) */
    double *edge_timestep;                  /* (3N) */
} oracle_state;

void oracle_rotate(double *q, double n1, double n2)
{
    double a = q[1], b = q[2];
    q[1] = n1 * a + n2 * b;
    q[2] = -n2 * a + n1 * b;
}

/* Central-upwind (Kurganov-Noelle-Petrova) edge flux in the edge-normal frame. */
void oracle_flux_function_central(const double *ql, const double *qr,
                                  double h_left, double h_right,
                                  double hle, double hre,
                                  double n1, double n2,
                                  double epsilon, double ze, double g,
                                  double *edgeflux, double *max_speed,
                                  double *pressure_flux, int64_t low_froude)
{
    double lrot[3], rrot[3], fl[3], fr[3];
    double ul, vl, uhl, vhl, ur, vr, uhr, vhr, inv;
    double cl, cr, s_max, s_min, denom, inv_den, local_fr;

    if (h_left == 0.0 && h_right == 0.0) {
        edgeflux[0] = edgeflux[1] = edgeflux[2] = 0.0;
        *max_speed = 0.0;
        *pressure_flux = 0.0;
        return;
    }
    lrot[0] = ql[0]; lrot[1] = ql[1]; lrot[2] = ql[2];
    rrot[0] = qr[0]; rrot[1] = qr[1]; rrot[2] = qr[2];
    oracle_rotate(lrot, n1, n2);
    oracle_rotate(rrot, n1, n2);

    uhl = lrot[1];
    vhl = lrot[2];
    if (hle > 0.0) {
        inv = 1.0 / hle;
        ul = uhl * inv;
        uhl = h_left * ul;
        vl = vhl * inv;
        vhl = h_left * inv * vhl;
    } else {
        ul = uhl = vhl = vl = 0.0;
    }
    uhr = rrot[1];
    vhr = rrot[2];
    if (hre > 0.0) {
        inv = 1.0 / hre;
        ur = uhr * inv;
        uhr = h_right * ur;
        vr = vhr * inv;
        vhr = h_right * inv * vhr;
    } else {
        ur = uhr = vhr = vr = 0.0;
    }

    cl = sqrt(g * h_left);
    cr = sqrt(g * h_right);

    if (low_froude == 1) {
        local_fr = sqrt(fmax(0.001, fmin(1.0,
            (ur * ur + ul * ul + vr * vr + vl * vl) /
            (cl * cl + cr * cr + 1.0e-10))));
    } else if (low_froude == 2) {
        local_fr = sqrt((ur * ur + ul * ul + vr * vr + vl * vl) /
                        (cl * cl + cr * cr + 1.0e-10));
        local_fr = sqrt(fmin(1.0, 0.01 + fmax(local_fr - 0.01, 0.0)));
    } else {
        local_fr = 1.0;
    }

    s_max = fmax(ul + cl, ur + cr);
    if (s_max < 0.0) s_max = 0.0;
    s_min = fmin(ul - cl, ur - cr);
    if (s_min > 0.0) s_min = 0.0;

    fl[0] = ul * h_left;  fl[1] = ul * uhl;  fl[2] = ul * vhl;
    fr[0] = ur * h_right; fr[1] = ur * uhr;  fr[2] = ur * vhr;

    denom = s_max - s_min;
    if (denom < epsilon) {
        edgeflux[0] = edgeflux[1] = edgeflux[2] = 0.0;
        *max_speed = 0.0;
        *pressure_flux = 0.5 * g * 0.5 * (h_left * h_left + h_right * h_right);
        return;
    }
    *max_speed = fmax(s_max, -s_min);
    inv_den = 1.0 / fmax(denom, 1.0e-100);
    edgeflux[0] = s_max * fl[0] - s_min * fr[0];
    edgeflux[0] += (s_max * s_min) * (fmax(rrot[0], ze) - fmax(lrot[0], ze));
    edgeflux[0] *= inv_den;
    edgeflux[1] = s_max * fl[1] - s_min * fr[1];
    edgeflux[1] += local_fr * (s_max * s_min) * (uhr - uhl);
    edgeflux[1] *= inv_den;
    edgeflux[2] = s_max * fl[2] - s_min * fr[2];
    edgeflux[2] += local_fr * (s_max * s_min) * (vhr - vhl);
    edgeflux[2] *= inv_den;
    *pressure_flux = 0.5 * g * (s_max * h_left * h_left - s_min * h_right * h_right) * inv_den;
    oracle_rotate(edgeflux, n1, -n2);
}

/*
 * Blend the edge flux towards the Villemonte-corrected broad-crested weir
 * discharge Q = Qfactor 2/3 H sqrt(2/3 g H) (1 - Q2/Q1)^0.385, back to the
 * shallow-water flux as the submergence ratio goes from s1 to s2 or the
 * tailwater head over the weir height from h1 to h2; momentum fluxes follow
 * the mass flux's scaling (capped at 10). Sets the edge's wave speed from the
 * weir head when the mass flux is nonzero.
 */
void oracle_adjust_edgeflux_with_weir(double *ef, double h_left, double h_right,
                                      double g, double weir_height, double Qfactor,
                                      double s1, double s2, double h1, double h2,
                                      double *max_speed)
{
    const double two3 = 2.0 / 3.0;
    double minhd = fmin(h_left, h_right);
    double maxhd = fmax(h_left, h_right);
    double q1 = Qfactor * two3 * maxhd * sqrt(two3 * g * maxhd);
    double q2 = Qfactor * two3 * minhd * sqrt(two3 * g * minhd);
    double q_ratio = q2 / fmax(q1, 1.0e-100);
    double h_ratio = minhd / fmax(maxhd, 1.0e-100);
    double head_ratio = minhd / fmax(weir_height, 1.0e-100);
    double q = q1 * oracle_pow(1.0 - q_ratio, 0.385);
    if (h_right > h_left) q *= -1.0;
    if ((h_ratio < s2) & (head_ratio < h2)) {
        double w1 = fmin(fmax(h_ratio - s1, 0.0) / (s2 - s1), 1.0);
        double w2 = fmin(fmax(head_ratio - h1, 0.0) / (h2 - h1), 1.0);
        double blended = (q * (1.0 - w1) + w1 * ef[0]) * (1.0 - w2) + w2 * ef[0];
        double scale = (fabs(ef[0]) > 1.0e-100) ? blended / ef[0] : 0.0;
        scale = fmax(scale, 0.0);
        ef[0] = blended;
        ef[1] *= fmin(scale, 10.0);
        ef[2] *= fmin(scale, 10.0);
    }
    if (fabs(ef[0]) > 0.0)
        *max_speed = sqrt(g * (maxhd + weir_height)) + fabs(ef[0] / (maxhd + 1.0e-12));
}

/*
 * The weir on riverwall rank r of edge (k, n): only where the wall stands above
 * both cells' beds and either side's first-order head over z_half is positive
 * (sw_domain_orig.c:905-968).
 */
static void riverwall_weir(const oracle_mesh *m, const oracle_params *p, const oracle_state *s,
                           int64_t k, int64_t n, int64_t r, double zl, double zr,
                           double zc, double zc_n, double hc_n, double z_half,
                           double *ef, double *speed)
{
    const double elev = m->riverwall_elevation[r];
    const double weir_height = fmax(elev - fmin(zl, zr), 0.0);
    const double *hp;
    double hl, hr;
    if (!(elev > fmax(zc, zc_n))) return;
    hl = fmax(s->w[k] - z_half, 0.0);
    hr = (n >= 0) ? fmax(s->w[n] - z_half, 0.0) : fmax(hc_n + zr - z_half, 0.0);
    if (!((hl > 0.0) || (hr > 0.0))) return;
    hp = m->riverwall_hydraulic_properties +
         m->riverwall_rowIndex[r] * m->ncol_riverwall_hydraulic_properties;
    oracle_adjust_edgeflux_with_weir(ef, hl, hr, p->g, weir_height, hp[0], hp[1], hp[2],
                                     hp[3], hp[4], speed);
}

/* Right-hand state of edge (k,i): neighbour's edge values or the boundary array. */
static void right_state(const oracle_mesh *m, const oracle_state *s, int64_t k,
                        int i, double zl, double *qr, double *zr, double *hre)
{
    int64_t ki = 3 * k + i;
    int64_t n = m->neighbours[ki];
    if (n < 0) {
        int64_t j = -n - 1;
        qr[0] = s->w_b[j];
        qr[1] = s->uh_b[j];
        qr[2] = s->vh_b[j];
        *zr = zl;
        *hre = fmax(qr[0] - *zr, 0.0);
    } else {
        int64_t nm = 3 * n + m->neighbour_edges[ki];
        qr[0] = s->w_e[nm];
        qr[1] = s->uh_e[nm];
        qr[2] = s->vh_e[nm];
        *zr = s->z_e[nm];
        *hre = s->h_e[nm];
    }
}

/*
 * compute_fluxes in the reference's structure: each interior edge is computed
 * once, from the lower-numbered triangle, and mirrored to the neighbour.
 * Returns the CFL timestep on substep 0, else `timestep`.
 * boundary_flux_sum[substep] accumulates the boundary inflow.
 */
/*
 * The edge loop and the sums of _compute_fluxes_central. An edge is skipped
 * when done[ki] == call (its mirror was set earlier in this call) or, with
 * update_next_flux given, when update_next_flux[ki] != 1 (:829); skipped
 * edges keep their stored edge_flux_work / pressuregrad_work / edge_timestep,
 * which the sums read like fresh ones. *local_timestep is the reference's
 * static local_timestep (:779): reset to 1e100 when `reset`, else lowered
 * from its last value (:813-817).
 */
static double compute_fluxes_core(const oracle_mesh *m, const oracle_params *p,
                                  oracle_state *s, int64_t substep, double timestep,
                                  double *boundary_flux_sum, double *edge_flux_work,
                                  double *pressuregrad_work, int64_t *done, int64_t call,
                                  const int64_t *update_next_flux, double *local_ts,
                                  int reset)
{
    const int64_t N = m->N;
    double local_timestep = reset ? 1.0e+100 : *local_ts;
    int64_t k, n, nm = 0, ki, rw_count = 0;
    int i;

    memset(s->ex_w, 0, N * sizeof(double));
    memset(s->ex_uh, 0, N * sizeof(double));
    memset(s->ex_vh, 0, N * sizeof(double));
    memset(s->max_speed, 0, N * sizeof(double));

    for (k = 0; k < N; k++) {
        for (i = 0; i < 3; i++) {
            double ql[3], qr[3], ef[3], zl, zr, hle, hre, hc, zc, hc_n, zc_n;
            double z_half, h_left, h_right, speed, pflux, length, bsw, tmp;
            int is_rw;
            ki = 3 * k + i;
            if (done[ki] == call || (update_next_flux && update_next_flux[ki] != 1)) {
                if (m->edge_flux_type && m->edge_flux_type[ki] == 1) rw_count++;
                continue;
            }
            ql[0] = s->w_e[ki];
            ql[1] = s->uh_e[ki];
            ql[2] = s->vh_e[ki];
            zl = s->z_e[ki];
            hc = s->h[k];
            zc = s->z[k];
            hle = s->h_e[ki];
            n = m->neighbours[ki];
            hc_n = hc;
            zc_n = s->z[k];
            right_state(m, s, k, i, zl, qr, &zr, &hre);
            if (n >= 0) {
                hc_n = s->h[n];
                zc_n = s->z[n];
                nm = 3 * n + m->neighbour_edges[ki];
            }
            z_half = fmax(zl, zr);
            is_rw = m->edge_flux_type && m->edge_flux_type[ki] == 1;
            if (is_rw) {
                rw_count++;
                z_half = fmax(m->riverwall_elevation[rw_count - 1], z_half);
            }
            h_left = fmax(hle + zl - z_half, 0.0);
            h_right = fmax(hre + zr - z_half, 0.0);
            oracle_flux_function_central(ql, qr, h_left, h_right, hle, hre,
                                         m->normals[2 * ki], m->normals[2 * ki + 1],
                                         p->epsilon, z_half, p->g, ef, &speed,
                                         &pflux, p->low_froude);
            if (is_rw)
                riverwall_weir(m, p, s, k, n, rw_count - 1, zl, zr, zc, zc_n, hc_n, z_half,
                               ef, &speed);
            length = m->edgelengths[ki];
            ef[0] *= length;
            ef[1] *= length;
            ef[2] *= length;
            edge_flux_work[3 * ki + 0] = -ef[0];
            edge_flux_work[3 * ki + 1] = -ef[1];
            edge_flux_work[3 * ki + 2] = -ef[2];
            bsw = length * (-p->g * 0.5 * (h_left * h_left - hle * hle - (hle + hc) * (zl - zc)) + pflux);
            pressuregrad_work[ki] = bsw;
            done[ki] = call;
            if (n >= 0) {
                edge_flux_work[3 * nm + 0] = ef[0];
                edge_flux_work[3 * nm + 1] = ef[1];
                edge_flux_work[3 * nm + 2] = ef[2];
                bsw = length * (-p->g * 0.5 * (h_right * h_right - hre * hre - (hre + hc_n) * (zr - zc_n)) + pflux);
                pressuregrad_work[nm] = bsw;
                done[nm] = call;
            }
            if (substep == 0) {
                tmp = 1.0 / fmax(speed, p->epsilon);
                s->edge_timestep[ki] = m->radii[k] * tmp;
                if (n >= 0) {
                    s->edge_timestep[nm] = m->radii[n] * tmp;
                    if (m->tri_full_flag[k] == 1)
                        s->max_speed[n] = fmax(s->max_speed[n], speed);
                }
                if (m->tri_full_flag[k] == 1) {
                    s->max_speed[k] = fmax(s->max_speed[k], speed);
                    if (speed > p->epsilon) {
                        local_timestep = fmin(local_timestep, s->edge_timestep[ki]);
                        if (n >= 0)
                            local_timestep = fmin(local_timestep, s->edge_timestep[nm]);
                    }
                }
            }
        }
    }

    for (k = 0; k < N; k++) {
        double inv_area;
        for (i = 0; i < 3; i++) {
            ki = 3 * k + i;
            n = m->neighbours[ki];
            s->ex_w[k] += edge_flux_work[3 * ki + 0];
            s->ex_uh[k] += edge_flux_work[3 * ki + 1];
            s->ex_vh[k] += edge_flux_work[3 * ki + 2];
            if (((n < 0) && (m->tri_full_flag[k] == 1)) ||
                ((n >= 0) && (m->tri_full_flag[k] == 1) && (m->tri_full_flag[n] == 0)))
                boundary_flux_sum[substep] += edge_flux_work[3 * ki];
            s->ex_uh[k] -= m->normals[6 * k + 2 * i] * pressuregrad_work[ki];
            s->ex_vh[k] -= m->normals[6 * k + 2 * i + 1] * pressuregrad_work[ki];
        }
        inv_area = 1.0 / m->areas[k];
        s->ex_w[k] *= inv_area;
        s->ex_uh[k] *= inv_area;
        s->ex_vh[k] *= inv_area;
    }
    *local_ts = local_timestep;
    if (substep == 0) timestep = local_timestep;
    return timestep;
}

double oracle_compute_fluxes(const oracle_mesh *m, const oracle_params *p,
                             oracle_state *s, int64_t substep, double timestep,
                             double *boundary_flux_sum, double *edge_flux_work,
                             double *pressuregrad_work, int64_t *done)
{
    double lt;
    memset(done, 0, 3 * m->N * sizeof(int64_t));
    return compute_fluxes_core(m, p, s, substep, timestep, boundary_flux_sum, edge_flux_work,
                               pressuregrad_work, done, 1, NULL, &lt, 1);
}

/*
 * Local flux updating (max_flux_update_frequency > 1; the reference allows
 * it with euler steps only, shallow_water_domain.py:1193-1215). lts holds the
 * reference's per-domain arrays and its static counters.
 */
typedef struct {
    int64_t max_flux_update_frequency;
    int64_t call;                    /* _compute_fluxes_central's static call */
    int64_t cyclic_number_of_steps;  /* _compute_flux_update_frequency's static, from -1 */
    double local_timestep;           /* _compute_fluxes_central's static local_timestep */
    int64_t allow_timestep_increase;
    int64_t *already_computed_flux;  /* (3N) */
    int64_t *flux_update_frequency;  /* (3N), from 1 */
    int64_t *update_next_flux;       /* (3N), from 1 */
    int64_t *update_extrapolation;   /* (N) */
    double *edge_flux_work;          /* (9N) */
    double *pressuregrad_work;       /* (3N) */
} oracle_lts;

/* compute_fluxes of a domain with local flux updating (substep 0 of euler):
 * call++ (sw_domain_orig.c:789), then the edge loop above with the skip test
 * and the static timestep reset only when allow_timestep_increase is 1. */
double oracle_compute_fluxes_lts(const oracle_mesh *m, const oracle_params *p,
                                 oracle_state *s, double timestep,
                                 double *boundary_flux_sum, oracle_lts *l)
{
    l->call++;
    return compute_fluxes_core(m, p, s, 0, timestep, boundary_flux_sum, l->edge_flux_work,
                               l->pressuregrad_work, l->already_computed_flux, l->call,
                               l->update_next_flux, &l->local_timestep,
                               l->allow_timestep_increase == 1);
}

static int64_t mod_pow2(int64_t n, int64_t d)
{
    return (int64_t)((uint64_t)n & ((uint64_t)d - 1u));
}

/* _compute_flux_update_frequency (sw_domain_orig.c:509-669), in its loop
 * order: part 1 re-derives each edge pair's frequency from its edge
 * timesteps after a flux update (the sign of already_computed_flux marks the
 * pair's first visit), part 2 takes the minimum over each triangle and then
 * over each edge pair and sets update_next_flux for the next step. */
void oracle_compute_flux_update_frequency(const oracle_mesh *m, const oracle_state *s,
                                          double timestep, oracle_lts *l)
{
    const int64_t N = m->N, M = l->max_flux_update_frequency;
    const double notSoFast = 1.0;
    int64_t k, i, k3, ki, mm, n, nm = 0, ii, ii2, fuf;
    int64_t *acf = l->already_computed_flux, *ff = l->flux_update_frequency;
    if (M == 1) return;
    l->cyclic_number_of_steps++;
    if (l->cyclic_number_of_steps == M) l->cyclic_number_of_steps = 0;
    for (k = 0; k < N; k++) {
        for (i = 0; i < 3; i++) {
            ki = 3 * k + i;
            if (mod_pow2(l->cyclic_number_of_steps, ff[ki]) != 0) continue;
            n = m->neighbours[ki];
            if (n >= 0) {
                mm = m->neighbour_edges[ki];
                nm = 3 * n + mm;
            }
            if (acf[ki] > 0) {
                acf[ki] *= -1;
                if (n >= 0) acf[nm] *= -1;
            } else {
                acf[ki] *= -1;
                if (n >= 0) acf[nm] *= -1;
                continue;
            }
            fuf = (int64_t)fmin((s->edge_timestep[ki] / timestep) * notSoFast, M * 1.);
            if (n >= 0)
                fuf = (int64_t)fmin((double)(int64_t)fmin(s->edge_timestep[nm] / timestep * notSoFast,
                                                          M * 1.),
                                    (double)fuf);
            if (fuf < 1) fuf = 1;
            if (fuf > M) fuf = M;
            ii = 2;
            while (ii < M) {
                ii2 = 2 * ii;
                if ((fuf > ii) && (fuf < ii2)) {
                    fuf = ii;
                    continue;
                }
                ii = ii2;
            }
            ff[ki] = fuf;
            if (n >= 0) ff[nm] = fuf;
        }
    }
    for (k = 0; k < N; k++) {
        k3 = 3 * k;
        ii = (int64_t)fmin((double)ff[k3], fmin((double)ff[k3 + 1], (double)ff[k3 + 2]));
        ff[k3] = (int64_t)fmin((double)ii, (double)ff[k3]);
        ff[k3 + 1] = (int64_t)fmin((double)ii, (double)ff[k3 + 1]);
        ff[k3 + 2] = (int64_t)fmin((double)ii, (double)ff[k3 + 2]);
    }
    for (k = 0; k < N; k++) {
        l->update_extrapolation[k] = 0;
        for (i = 0; i < 3; i++) {
            ki = 3 * k + i;
            n = m->neighbours[ki];
            if (n >= 0) {
                nm = 3 * n + m->neighbour_edges[ki];
                ff[ki] = (int64_t)fmin((double)ff[ki], (double)ff[nm]);
            }
            if (mod_pow2(l->cyclic_number_of_steps + 1, ff[ki]) == 0) {
                l->update_next_flux[ki] = 1;
                l->update_extrapolation[k] = 1;
            } else {
                l->update_next_flux[ki] = 0;
            }
        }
    }
    l->allow_timestep_increase = (l->cyclic_number_of_steps + 1 == M) ? 1 : 0;
}

/*
 * Same fluxes with each triangle evaluating its own three edges (the mirrored
 * value of an interior edge is recomputed from the neighbour's side, which is
 * bit-identical for low_froude 0). Triangles are independent, so the loop runs
 * OpenMP-parallel with a min-reduction for the CFL timestep; per-triangle
 * boundary inflow goes to bflux_tri (N) and is summed serially.
 */
double oracle_compute_fluxes_gather(const oracle_mesh *m, const oracle_params *p,
                                    oracle_state *s, int64_t substep, double timestep,
                                    double *boundary_flux_sum, double *bflux_tri)
{
    const int64_t N = m->N;
    double local_timestep = 1.0e+100;
    int64_t k;

#pragma omp parallel for schedule(static) reduction(min:local_timestep)
    for (k = 0; k < N; k++) {
        double sw = 0.0, sx = 0.0, sy = 0.0, bf = 0.0, inv_area;
        double hc = s->h[k], zc = s->z[k];
        int i;
        for (i = 0; i < 3; i++) {
            int64_t ki = 3 * k + i;
            int64_t n = m->neighbours[ki];
            double ql[3], qr[3], ef[3], zl, zr, hle, hre, z_half, h_left, h_right;
            double speed, pflux, length, pg, tmp;
            int64_t rw = m->edge_flux_type ? m->edge_river_wall_counter[ki] - 1 : -1;
            ql[0] = s->w_e[ki];
            ql[1] = s->uh_e[ki];
            ql[2] = s->vh_e[ki];
            zl = s->z_e[ki];
            hle = s->h_e[ki];
            right_state(m, s, k, i, zl, qr, &zr, &hre);
            z_half = fmax(zl, zr);
            if (rw >= 0) z_half = fmax(m->riverwall_elevation[rw], z_half);
            h_left = fmax(hle + zl - z_half, 0.0);
            h_right = fmax(hre + zr - z_half, 0.0);
            oracle_flux_function_central(ql, qr, h_left, h_right, hle, hre,
                                         m->normals[2 * ki], m->normals[2 * ki + 1],
                                         p->epsilon, z_half, p->g, ef, &speed,
                                         &pflux, p->low_froude);
            if (rw >= 0)
                riverwall_weir(m, p, s, k, n, rw, zl, zr, zc, (n >= 0) ? s->z[n] : zc,
                               (n >= 0) ? s->h[n] : hc, z_half, ef, &speed);
            length = m->edgelengths[ki];
            ef[0] *= length;
            ef[1] *= length;
            ef[2] *= length;
            pg = length * (-p->g * 0.5 * (h_left * h_left - hle * hle - (hle + hc) * (zl - zc)) + pflux);
            sw += -ef[0];
            sx += -ef[1];
            sy += -ef[2];
            if (((n < 0) && (m->tri_full_flag[k] == 1)) ||
                ((n >= 0) && (m->tri_full_flag[k] == 1) && (m->tri_full_flag[n] == 0)))
                bf += -ef[0];
            sx -= m->normals[2 * ki] * pg;
            sy -= m->normals[2 * ki + 1] * pg;
            if (substep == 0 && m->tri_full_flag[k] == 1 && speed > p->epsilon) {
                tmp = 1.0 / fmax(speed, p->epsilon);
                local_timestep = fmin(local_timestep, m->radii[k] * tmp);
            }
        }
        inv_area = 1.0 / m->areas[k];
        s->ex_w[k] = sw * inv_area;
        s->ex_uh[k] = sx * inv_area;
        s->ex_vh[k] = sy * inv_area;
        bflux_tri[k] = bf;
    }
    for (k = 0; k < N; k++) boundary_flux_sum[substep] += bflux_tri[k];
    if (substep == 0) timestep = local_timestep;
    return timestep;
}

double oracle_protect(const oracle_mesh *m, const oracle_params *p, oracle_state *s)
{
    double mass_error = 0.0;
    int64_t k;
#pragma omp parallel for schedule(static) reduction(+:mass_error)
    for (k = 0; k < m->N; k++) {
        double hc = s->w[k] - s->z[k];
        if (hc < p->minimum_allowed_height * 1.0) {
            s->uh[k] = 0.0;
            s->vh[k] = 0.0;
            if (hc <= 0.0) {
                double bmin = s->z[k];
                if (s->w[k] < bmin) {
                    mass_error += (bmin - s->w[k]) * m->areas[k];
                    s->w[k] = bmin;
                    s->w_v[3 * k] = bmin;
                    s->w_v[3 * k + 1] = bmin;
                    s->w_v[3 * k + 2] = bmin;
                }
            }
        }
    }
    return mass_error;
}

static void qmin_qmax(double dq0, double dq1, double dq2, double *qmin, double *qmax)
{
    *qmax = fmax(fmax(dq0, fmax(dq0 + dq1, dq0 + dq2)), 0.0);
    *qmin = fmin(fmin(dq0, fmin(dq0 + dq1, dq0 + dq2)), 0.0);
}

static void limit(double *dqv, double qmin, double qmax, double beta)
{
    double r = 1000.0, r0 = 1.0, phi;
    const double tiny = 1.0e-100;
    int i;
    for (i = 0; i < 3; i++) {
        if (dqv[i] < -tiny) r0 = qmin / dqv[i];
        if (dqv[i] > tiny) r0 = qmax / dqv[i];
        r = fmin(r0, r);
    }
    phi = fmin(r * beta, 1.0);
    dqv[0] = dqv[0] * phi;
    dqv[1] = dqv[1] * phi;
    dqv[2] = dqv[2] * phi;
}

/* Limited jumps of one quantity from the auxiliary-triangle gradient. */
static void aux_jumps(const double *q, int64_t k, int64_t k0, int64_t k1, int64_t k2,
                      double dx1, double dx2, double dy1, double dy2, double inv_area2,
                      double dxv0, double dxv1, double dxv2,
                      double dyv0, double dyv1, double dyv2, double beta,
                      double *out)
{
    double dq0 = q[k0] - q[k];
    double dq1 = q[k1] - q[k0];
    double dq2 = q[k2] - q[k0];
    double a = dy2 * dq1 - dy1 * dq2;
    double b, qmin, qmax, dqv[3];
    a *= inv_area2;
    b = dx1 * dq2 - dx2 * dq1;
    b *= inv_area2;
    dqv[0] = a * dxv0 + b * dyv0;
    dqv[1] = a * dxv1 + b * dyv1;
    dqv[2] = a * dxv2 + b * dyv2;
    qmin_qmax(dq0, dq1, dq2, &qmin, &qmax);
    limit(dqv, qmin, qmax, beta);
    out[0] = q[k] + dqv[0];
    out[1] = q[k] + dqv[1];
    out[2] = q[k] + dqv[2];
}

/* Limited jumps along the direction to the single interior neighbour. */
static void line_jumps(const double *q, int64_t k, int64_t k1, double dx2, double dy2,
                       double dxv0, double dxv1, double dxv2,
                       double dyv0, double dyv1, double dyv2, double beta, double *out)
{
    double dq1 = q[k1] - q[k];
    double a = dq1 * dx2, b = dq1 * dy2, qmin, qmax, dqv[3];
    dqv[0] = a * dxv0 + b * dyv0;
    dqv[1] = a * dxv1 + b * dyv1;
    dqv[2] = a * dxv2 + b * dyv2;
    if (dq1 >= 0.0) { qmin = 0.0; qmax = dq1; } else { qmin = dq1; qmax = 0.0; }
    limit(dqv, qmin, qmax, beta);
    out[0] = q[k] + dqv[0];
    out[1] = q[k] + dqv[1];
    out[2] = q[k] + dqv[2];
}

static void const_edges(oracle_state *s, int64_t k)
{
    int i;
    for (i = 0; i < 3; i++) {
        s->w_e[3 * k + i] = s->w[k];
        s->uh_e[3 * k + i] = s->uh[k];
        s->vh_e[3 * k + i] = s->vh[k];
        s->h_e[3 * k + i] = s->h[k];
    }
}

/* Returns 0, or -1 when a two-boundary triangle has no interior neighbour. */
int64_t oracle_extrapolate(const oracle_mesh *m, const oracle_params *p, oracle_state *s)
{
    const int64_t N = m->N;
    const double mah = p->minimum_allowed_height;
    const double a_tmp = 0.3, b_tmp = 0.1;
    const double c_tmp = 1.0 / (a_tmp - b_tmp);
    const double d_tmp = 1.0 - (c_tmp * a_tmp);
    int64_t k;

    memset(s->xwork, 0, N * sizeof(double));
    memset(s->ywork, 0, N * sizeof(double));

    if (p->extrapolate_velocity_second_order == 1) {
#pragma omp parallel for schedule(static)
        for (k = 0; k < N; k++) {
            double dk;
            s->h[k] = fmax(s->w[k] - s->z[k], 0.0);
            dk = s->h[k];
            if (dk > mah) {
                double inv = 1.0 / dk;
                s->xwork[k] = s->uh[k];
                s->uh[k] = s->uh[k] * inv;
                s->ywork[k] = s->vh[k];
                s->vh[k] = s->vh[k] * inv;
            } else {
                s->xwork[k] = 0.0;
                s->uh[k] = 0.0;
                s->ywork[k] = 0.0;
                s->vh[k] = 0.0;
            }
        }
    }

#pragma omp parallel for schedule(static)
    for (k = 0; k < N; k++) {
        int64_t k0 = m->surrogate[3 * k], k1 = m->surrogate[3 * k + 1], k2 = m->surrogate[3 * k + 2];
        if (((s->h[k0] < mah) | (k0 == k)) & ((s->h[k1] < mah) | (k1 == k)) &
            ((s->h[k2] < mah) | (k2 == k))) {
            s->xwork[k] = 0.0;
            s->uh[k] = 0.0;
            s->ywork[k] = 0.0;
            s->vh[k] = 0.0;
        }
    }

    int64_t status = 0;
#pragma omp parallel for schedule(static) reduction(min:status)
    for (k = 0; k < N; k++) {
        int64_t k3 = 3 * k, k6 = 6 * k;
        double x, y, dxv0, dxv1, dxv2, dyv0, dyv1, dyv2;
        if (m->nboundaries[k] == 3) {
            const_edges(s, k);
            continue;
        }
        x = m->centroids[2 * k];
        y = m->centroids[2 * k + 1];
        dxv0 = m->edge_midpoints[k6] - x;
        dxv1 = m->edge_midpoints[k6 + 2] - x;
        dxv2 = m->edge_midpoints[k6 + 4] - x;
        dyv0 = m->edge_midpoints[k6 + 1] - y;
        dyv1 = m->edge_midpoints[k6 + 3] - y;
        dyv2 = m->edge_midpoints[k6 + 5] - y;

        if (m->nboundaries[k] <= 1) {
            int64_t k0 = m->surrogate[k3], k1 = m->surrogate[k3 + 1], k2 = m->surrogate[k3 + 2];
            double x0 = m->centroids[2 * k0], y0 = m->centroids[2 * k0 + 1];
            double x1 = m->centroids[2 * k1], y1 = m->centroids[2 * k1 + 1];
            double x2 = m->centroids[2 * k2], y2 = m->centroids[2 * k2 + 1];
            double dx1 = x1 - x0, dx2 = x2 - x0, dy1 = y1 - y0, dy2 = y2 - y0;
            double area2 = dy2 * dx1 - dy1 * dx2;
            double hc, h0, h1, h2, hmin, hmax, hfactor, inv_area2, beta;
            if (area2 <= 0.0) {
                const_edges(s, k);
                continue;
            }
            hc = s->h[k]; h0 = s->h[k0]; h1 = s->h[k1]; h2 = s->h[k2];
            hmin = fmin(fmin(h0, fmin(h1, h2)), hc);
            hmax = fmax(fmax(h0, fmax(h1, h2)), hc);
            hfactor = fmax(0.0, fmin(c_tmp * fmax(hmin, 0.0) / fmax(hc, 1.0e-06) + d_tmp,
                                     fmin(c_tmp * fmax(hc, 0.0) / fmax(hmax, 1.0e-06) + d_tmp, 1.0)));
            hfactor = fmin(1.2 * fmax(hmin - mah, 0.0) / (fmax(hmin, 0.0) + 1.0 * mah), hfactor);
            inv_area2 = 1.0 / area2;

            beta = p->beta_w_dry + (p->beta_w - p->beta_w_dry) * hfactor;
            if (beta > 0.0) {
                aux_jumps(s->w, k, k0, k1, k2, dx1, dx2, dy1, dy2, inv_area2,
                          dxv0, dxv1, dxv2, dyv0, dyv1, dyv2, beta, &s->w_e[k3]);
                aux_jumps(s->h, k, k0, k1, k2, dx1, dx2, dy1, dy2, inv_area2,
                          dxv0, dxv1, dxv2, dyv0, dyv1, dyv2, beta, &s->h_e[k3]);
            } else {
                s->w_e[k3] = s->w_e[k3 + 1] = s->w_e[k3 + 2] = s->w[k];
                s->h_e[k3] = s->h_e[k3 + 1] = s->h_e[k3 + 2] = s->h[k];
            }
            beta = p->beta_uh_dry + (p->beta_uh - p->beta_uh_dry) * hfactor;
            if (beta > 0.0) {
                aux_jumps(s->uh, k, k0, k1, k2, dx1, dx2, dy1, dy2, inv_area2,
                          dxv0, dxv1, dxv2, dyv0, dyv1, dyv2, beta, &s->uh_e[k3]);
            } else {
                s->uh_e[k3] = s->uh_e[k3 + 1] = s->uh_e[k3 + 2] = s->uh[k];
            }
            beta = p->beta_vh_dry + (p->beta_vh - p->beta_vh_dry) * hfactor;
            if (beta > 0.0) {
                aux_jumps(s->vh, k, k0, k1, k2, dx1, dx2, dy1, dy2, inv_area2,
                          dxv0, dxv1, dxv2, dyv0, dyv1, dyv2, beta, &s->vh_e[k3]);
            } else {
                s->vh_e[k3] = s->vh_e[k3 + 1] = s->vh_e[k3 + 2] = s->vh[k];
            }
        } else {
            /* two boundaries: gradient along the single interior neighbour */
            int64_t e, k1;
            double dx1, dy1, area2, dx2, dy2;
            for (e = k3; e < k3 + 3; e++)
                if (m->surrogate[e] != k) break;
            if (e == k3 + 3) {
                status = -1;
                continue;
            }
            k1 = m->surrogate[e];
            dx1 = m->centroids[2 * k1] - x;
            dy1 = m->centroids[2 * k1 + 1] - y;
            area2 = dx1 * dx1 + dy1 * dy1;
            dx2 = 1.0 / area2;
            dy2 = dx2 * dy1;
            dx2 *= dx1;
            line_jumps(s->w, k, k1, dx2, dy2, dxv0, dxv1, dxv2, dyv0, dyv1, dyv2, p->beta_w, &s->w_e[k3]);
            line_jumps(s->h, k, k1, dx2, dy2, dxv0, dxv1, dxv2, dyv0, dyv1, dyv2, p->beta_w, &s->h_e[k3]);
            line_jumps(s->uh, k, k1, dx2, dy2, dxv0, dxv1, dxv2, dyv0, dyv1, dyv2, p->beta_w, &s->uh_e[k3]);
            line_jumps(s->vh, k, k1, dx2, dy2, dxv0, dxv1, dxv2, dyv0, dyv1, dyv2, p->beta_w, &s->vh_e[k3]);
        }
    }

    if (status != 0) return status;
#pragma omp parallel for schedule(static)
    for (k = 0; k < N; k++) {
        int64_t k3 = 3 * k;
        int i;
        double *e;
        if (p->extrapolate_velocity_second_order == 1) {
            s->uh[k] = s->xwork[k];
            s->vh[k] = s->ywork[k];
        }
        e = s->w_e;
        s->w_v[k3] = e[k3 + 1] + e[k3 + 2] - e[k3];
        s->w_v[k3 + 1] = e[k3] + e[k3 + 2] - e[k3 + 1];
        s->w_v[k3 + 2] = e[k3] + e[k3 + 1] - e[k3 + 2];
        e = s->h_e;
        s->h_v[k3] = e[k3 + 1] + e[k3 + 2] - e[k3];
        s->h_v[k3 + 1] = e[k3] + e[k3 + 2] - e[k3 + 1];
        s->h_v[k3 + 2] = e[k3] + e[k3 + 1] - e[k3 + 2];
        if (p->extrapolate_velocity_second_order == 1) {
            for (i = 0; i < 3; i++) {
                double dk = s->h_e[k3 + i];
                s->uh_e[k3 + i] = s->uh_e[k3 + i] * dk;
                s->vh_e[k3 + i] = s->vh_e[k3 + i] * dk;
            }
        }
        e = s->uh_e;
        s->uh_v[k3] = e[k3 + 1] + e[k3 + 2] - e[k3];
        s->uh_v[k3 + 1] = e[k3] + e[k3 + 2] - e[k3 + 1];
        s->uh_v[k3 + 2] = e[k3] + e[k3 + 1] - e[k3 + 2];
        e = s->vh_e;
        s->vh_v[k3] = e[k3 + 1] + e[k3 + 2] - e[k3];
        s->vh_v[k3 + 1] = e[k3] + e[k3 + 2] - e[k3 + 1];
        s->vh_v[k3 + 2] = e[k3] + e[k3 + 1] - e[k3 + 2];
        for (i = 0; i < 3; i++) s->z_e[k3 + i] = s->w_e[k3 + i] - s->h_e[k3 + i];
        e = s->z_e;
        s->z_v[k3] = e[k3 + 1] + e[k3 + 2] - e[k3];
        s->z_v[k3 + 1] = e[k3] + e[k3 + 2] - e[k3 + 1];
        s->z_v[k3 + 2] = e[k3] + e[k3 + 1] - e[k3 + 2];
    }
    return 0;
}

void oracle_manning_friction_flat(int64_t N, double g, double eps,
                                  const double *w, const double *z,
                                  const double *uh, const double *vh,
                                  const double *eta, double *xmom_si, double *ymom_si)
{
    const double seven_thirds = 7.0 / 3.0;
    int64_t k;
#pragma omp parallel for schedule(static)
    for (k = 0; k < N; k++) {
        if (eta[k] > eps) {
            double h = w[k] - z[k];
            if (h >= eps) {
                double S = -g * eta[k] * eta[k] * sqrt((uh[k] * uh[k] + vh[k] * vh[k]));
                S /= pow(h, seven_thirds);
                xmom_si[k] += S * uh[k];
                ymom_si[k] += S * vh[k];
            }
        }
    }
}

/* Manning friction with the bed slope factor sqrt(1 + zx^2 + zy^2) of the
 * plane through the vertex beds and the depth over their mean. */
void oracle_manning_friction_sloped(int64_t N, double g, double eps, const double *x,
                                    const double *w, const double *zv,
                                    const double *uh, const double *vh,
                                    const double *eta, double *xmom_si, double *ymom_si)
{
    const double one_third = 1.0 / 3.0;
    const double seven_thirds = 7.0 / 3.0;
    int64_t k;
#pragma omp parallel for schedule(static)
    for (k = 0; k < N; k++) {
        if (eta[k] > eps) {
            const double *v = x + 6 * k;
            double z0 = zv[3 * k], z1 = zv[3 * k + 1], z2 = zv[3 * k + 2];
            double det = (v[5] - v[1]) * (v[2] - v[0]) - (v[3] - v[1]) * (v[4] - v[0]);
            double zx = (v[5] - v[1]) * (z1 - z0) - (v[3] - v[1]) * (z2 - z0);
            double zy = (v[2] - v[0]) * (z2 - z0) - (v[4] - v[0]) * (z1 - z0);
            double zs, h;
            zx /= det;
            zy /= det;
            zs = sqrt(1.0 + zx * zx + zy * zy);
            h = w[k] - (z0 + z1 + z2) * one_third;
            if (h >= eps) {
                double S = -g * eta[k] * eta[k] * zs * sqrt((uh[k] * uh[k] + vh[k] * vh[k]));
                S /= pow(h, seven_thirds);
                xmom_si[k] += S * uh[k];
                ymom_si[k] += S * vh[k];
            }
        }
    }
}

/* Returns 0, or -1 on a non-positive semi-implicit denominator. */
int64_t oracle_update(int64_t N, double dt, double *q, const double *ex, double *si)
{
    int64_t k, status = 0;
#pragma omp parallel for schedule(static) reduction(min:status)
    for (k = 0; k < N; k++) {
        double x = q[k], den;
        if (x == 0.0) si[k] = 0.0;
        else si[k] /= x;
        q[k] += dt * ex[k];
        den = 1.0 - dt * si[k];
        if (den <= 0.0) status = -1;
        else q[k] /= den;
    }
    if (status != 0) return -1;
    memset(si, 0, N * sizeof(double));
    return 0;
}

int64_t oracle_fix_negative_cells(int64_t N, const int64_t *tri_full_flag,
                                  double *w, const double *z, double *uh, double *vh)
{
    int64_t k, count = 0;
#pragma omp parallel for schedule(static) reduction(+:count)
    for (k = 0; k < N; k++) {
        if ((w[k] - z[k] < 0.0) & (tri_full_flag[k] > 0)) {
            count++;
            w[k] = z[k];
            uh[k] = 0.0;
            vh[k] = 0.0;
        }
    }
    return count;
}

void oracle_backup(int64_t N, const double *q, double *backup)
{
    memcpy(backup, q, N * sizeof(double));
}

void oracle_saxpy(int64_t N, double a, double b, double *q, const double *backup)
{
    int64_t k;
#pragma omp parallel for schedule(static)
    for (k = 0; k < N; k++) q[k] = a * q[k] + b * backup[k];
}

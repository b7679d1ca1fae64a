/*
 * anuga_hip.h - C ABI of the MI355X (gfx950) DE shallow-water step.
 *
 * One `anuga_domain` handle owns every device buffer of one mesh: the conserved
 * state (stage, xmomentum, ymomentum), bed elevation, friction and the static
 * mesh arrays stay resident in HBM for the whole evolve loop. Host pointers are
 * borrowed only for the duration of a call. Arrays crossing this boundary use
 * the reference's layout and triangle order (FP64, int64 indices, (N,3)
 * row-major); any internal renumbering is hidden behind it.
 *
 * Every entry point returns ANUGA_OK (0) or a negative ANUGA_E* code; the
 * message of the last failure is available from anuga_last_error().
 * Handles are independent; one handle must not be used from two threads at once.
 *
 * The reference binds the same work per call through Cython shims that rebuild
 * a `struct domain` of borrowed numpy pointers on every call
 * (anuga/shallow_water/sw_domain_orig_ext.pyx:104-355); the entry points below
 * replace those functions as cited on each one.
 */
#ifndef ANUGA_HIP_H
#define ANUGA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ANUGA_OK 0
#define ANUGA_E_ARG (-1)             /* invalid argument or unsupported option */
#define ANUGA_E_HIP (-2)             /* HIP runtime failure (allocation, launch) */
#define ANUGA_E_SEMI_IMPLICIT (-3)   /* semi-implicit denominator <= 0 (quantity.c:805-808) */
#define ANUGA_E_NO_NEIGHBOUR (-4)    /* two-boundary triangle without an interior neighbour
                                        (sw_domain_orig.c:2432-2437) */
#define ANUGA_E_STATE (-5)           /* call out of sequence */
#define ANUGA_E_TIMESTEP (-6)        /* too small timestep after the first-order fallback,
                                        or time overshot finaltime (generic_domain.py:2370-2391,
                                        :1873-1878) */

/* quantity ids */
#define ANUGA_STAGE 0
#define ANUGA_XMOMENTUM 1
#define ANUGA_YMOMENTUM 2
#define ANUGA_ELEVATION 3
#define ANUGA_FRICTION 4
#define ANUGA_HEIGHT 5

/* boundary kinds, one per boundary edge j (neighbours = -(j+1)) */
#define ANUGA_BC_REFLECTIVE 0         /* anuga/shallow_water/boundaries.py:235-292 */
#define ANUGA_BC_TRANSMISSIVE 1       /* generic_boundary_conditions.py:173-195 (edge values) */
#define ANUGA_BC_DIRICHLET 2          /* generic_boundary_conditions.py:221-264 */
#define ANUGA_BC_TRANSMISSIVE_CENTROID 3 /* transmissive with centroid_transmissive_bc */
#define ANUGA_BC_SET_STAGE_NORMAL_MOMENTUM 4   /* stage from the values, normal momentum kept,
                                                  tangential zero (boundaries.py:384-517) */
#define ANUGA_BC_TRANSMISSIVE_STAGE_ZERO_MOMENTUM 5 /* boundaries.py:520-550 */
#define ANUGA_BC_SET_STAGE_TRANSMISSIVE_MOMENTUM 6 /* stage from the values, momentum kept
                                                     (boundaries.py:295-381) */

/* arithmetic of the step */
#define ANUGA_PRECISION_FP64 0   /* everything FP64, bit-comparable with the reference */
#define ANUGA_PRECISION_MIXED 1  /* FP32 extrapolation/reconstruction/flux; FP64 state,
                                    update, CFL reduction, wet/dry thresholds, volume */

typedef struct anuga_domain anuga_domain;

/* Static mesh arrays in the reference's layout (anuga/abstract_2d_finite_volumes/
 * general_mesh.py, neighbour_mesh.py). Boundary edges hold neighbours = -(j+1). */
typedef struct anuga_mesh {
    int64_t number_of_elements;            /* N */
    int64_t boundary_length;               /* M */
    const int64_t *neighbours;             /* (N,3) */
    const int64_t *neighbour_edges;        /* (N,3) */
    const int64_t *surrogate_neighbours;   /* (N,3) */
    const int64_t *number_of_boundaries;   /* (N) */
    const double *normals;                 /* (N,6) */
    const double *edgelengths;             /* (N,3) */
    const double *radii;                   /* (N) */
    const double *areas;                   /* (N) */
    const double *centroid_coordinates;    /* (N,2) */
    const double *edge_coordinates;        /* (N,6) edge midpoints */
    const double *vertex_coordinates;      /* (N,6), or NULL; needed by sloped friction */
} anuga_mesh;

/* Algorithm parameters; the DE presets are in shallow_water_domain.py:583-707. */
typedef struct anuga_params {
    double g;
    double epsilon;
    double minimum_allowed_height;
    double beta_w, beta_w_dry;
    double beta_uh, beta_uh_dry;
    double beta_vh, beta_vh_dry;
    int64_t low_froude;                        /* 0 (the DE presets), 1 or 2 (FP64 path) */
    int64_t extrapolate_velocity_second_order; /* only 1 (the DE presets) */
    int64_t timestep_fluxcalls;                /* 1 = euler (DE0), 2 = rk2 (DE1), 3 = rk3 (DE2) */
    int64_t precision;                         /* ANUGA_PRECISION_* */
    int64_t reorder;                           /* internal triangle order: 0 = the caller's,
                                                  1 = Reverse-Cuthill-McKee, 2 = compact
                                                  tiles (one FP64 scenario: the fused
                                                  extrapolation + flux kernel; results are
                                                  the same bits in every order) */
    int64_t use_sloped_mannings;               /* 1 = _manning_friction_sloped (FP64 path;
                                                  needs vertex_coordinates) */
} anuga_params;

/* Create a handle: uploads the mesh, allocates all device buffers.
 * Replaces the per-call marshalling of get_python_domain_pointers
 * (sw_domain_orig_ext.pyx:143-355) with a one-off upload. */
int anuga_create(const anuga_mesh *mesh, const anuga_params *params, anuga_domain **out);
void anuga_destroy(anuga_domain *d);
/* Message of the last failure on this handle (or of the last failed create when d is NULL). */
const char *anuga_last_error(const anuga_domain *d);
/* Change algorithm parameters (flow algorithm presets, betas) between steps. */
int anuga_set_params(anuga_domain *d, const anuga_params *params);

/* Centroid values in the caller's triangle order (Quantity.centroid_values). */
int anuga_set_centroid_values(anuga_domain *d, int quantity, const double *values);
int anuga_get_centroid_values(anuga_domain *d, int quantity, double *values);
/* Edge values (N,3) from the last extrapolation; quantity in
 * STAGE, XMOMENTUM, YMOMENTUM, ELEVATION, HEIGHT (Quantity.edge_values).
 * After device steps that ran the fused kernel (reorder 2), which keeps edge
 * values on chip, this extrapolates the current state first. */
int anuga_get_edge_values(anuga_domain *d, int quantity, double *values);

/* Boundary condition of every boundary edge: kind (M) and, for Dirichlet edges,
 * [stage, xmom, ymom] (M,3) (may be NULL when no edge is Dirichlet).
 * Replaces Generic_Domain.update_boundary (generic_domain.py:2288-2306): the
 * boundary state is evaluated on the device inside the flux kernel. */
int anuga_set_boundary(anuga_domain *d, const int32_t *kind, const double *dirichlet_values);
/* Time-dependent boundary values (M,3) for one substep time: slot 0 at the
 * step's start t, slot 1 at t + dt (the second substep of rk2 and rk3), slot
 * 2 at t + dt/2 (rk3's third), the times at which the reference's
 * update_boundary evaluates Time_boundary & co. (generic_domain.py:1974-2179,
 * generic_boundary_conditions.py:297-417). Values are read by the kinds that
 * take them (DIRICHLET, SET_STAGE_*). Once a slot above 0 is set, each
 * substep reads its own slot, and device-controlled step runs
 * (anuga_evolve_steps) are refused until anuga_set_boundary is called again. */
int anuga_set_boundary_values(anuga_domain *d, int64_t slot, const double *values);

/* Riverwalls (FP64 path), in the layout the reference's flux loop reads
 * (struct domain fields edge_flux_type, riverwall_elevation,
 * riverwall_rowIndex, riverwall_hydraulic_properties, sw_domain.h, filled by
 * RiverWall.create_riverwalls, anuga/structures/riverwall.py:382-469):
 * edge_flux_type (N,3) in the caller's order, 1 on a wall half-edge; the
 * elevation and hydraulic-table row of each of the number_of_riverwall_edges
 * flagged half-edges in (triangle, edge) order; the table (nrows, ncols >= 5)
 * of Qfactor, s1, s2, h1, h2. Replaces the riverwall branch of
 * _compute_fluxes_central (sw_domain_orig.c:828-837, :880-968) and
 * adjust_edgeflux_with_weir (:671-763); an interior edge is a wall iff the
 * half-edge of its lower-numbered triangle is flagged, as in that loop.
 * number_of_riverwall_edges 0 removes the walls. */
int anuga_set_riverwalls(anuga_domain *d, const int64_t *edge_flux_type,
                         int64_t number_of_riverwall_edges, const double *riverwall_elevation,
                         const int64_t *riverwall_rowIndex, int64_t nrows, int64_t ncols,
                         const double *hydraulic_properties);

/* ---- Fractional-step operators ---------------------------------------------------
 * Rate_operator (anuga/operators/rate_operators.py:149-260): after every step
 * (apply_fractional_steps, generic_domain.py:1849) the stage of the
 * operator's triangles (indices in the caller's order, or all when NULL)
 * rises by factor * dt * rate, with rates[i] per triangle or the scalar
 * `rate` when rates is NULL; when a rate is negative the loss is capped at
 * the depth and the momenta scale with the remaining depth. Device-controlled
 * steps apply the operators inside the step graph; host-sequenced steps call
 * anuga_apply_operators after each step. *id receives the operator's number. */
int anuga_add_rate_operator(anuga_domain *d, int64_t n, const int64_t *indices,
                            const double *rates, double rate, double factor, int64_t *id);
/* New rates / factor for operator id (e.g. a rate function evaluated at a new time). */
int anuga_update_rate_operator(anuga_domain *d, int64_t id, const double *rates, double rate,
                               double factor);

/* Set_stage_operator (anuga/operators/set_stage_operator.py:25-58, applied
 * by Set_quantity.__call__, anuga/operators/set_quantity.py:95-131): after
 * every step the stage centroid values of the n triangles (caller's order;
 * indices NULL: all triangles) become stages[i], or stage when stages is
 * NULL. Applied in the order operators were added, rate operators included;
 * update with anuga_update_rate_operator (rates = stages, factor unused). */
int anuga_add_set_stage_operator(anuga_domain *d, int64_t n, const int64_t *indices,
                                 const double *stages, double stage, int64_t *id);
int anuga_apply_operators(anuga_domain *d);

/* Inlet_operator (anuga/structures/inlet_operator.py:75-183) on the device: a
 * discharge 0.5 (Q1 + Q2) over each step into n <= 2048 triangles (indices in
 * the caller's numbering, in the operator's order; areas and
 * total_area = np.sum(areas) in that order). A positive volume fills the
 * lowest stages first to a level surface (Inlet.set_stages_evenly,
 * anuga/structures/inlet.py:192-225; stages sorted ascending, equal stages by
 * position), a negative one lowers the region to a uniform depth or empties
 * it; momenta from the region's velocities (velocity_mode 0), from (u, v)
 * (1) or zero (2). The step's dt is the device's. Applied after every step,
 * with the other operators, in the order added. */
int anuga_add_inlet_operator(anuga_domain *d, int64_t n, const int64_t *indices,
                             const double *areas, double total_area, int64_t velocity_mode,
                             double u, double v, double Q1, double Q2, int64_t *id);
/* Q at the operator's time and one timestep later, for the next application. */
int anuga_update_inlet_operator(anuga_domain *d, int64_t id, double Q1, double Q2);
/* stats[4]: cumulative volume (fractional_step_volume_integral's share), total
 * applied volume, total requested volume, last applied discharge. Synchronizes. */
int anuga_get_inlet_operator_stats(anuga_domain *d, int64_t id, double *stats);
/* Apply one operator (host-sequenced steps that interleave host-side operators). */
int anuga_apply_operator(anuga_domain *d, int64_t id);

/* State of a subset of triangles (caller's indices): get gives (n,4)
 * {stage, xmomentum, ymomentum, elevation}; set takes (n,4) and writes
 * stage, xmomentum, ymomentum (indices distinct). For host-side operators on
 * small regions (Inlet_operator, anuga/structures/inlet_operator.py:75-165,
 * through Inlet's take/put of centroid values, inlet.py:63-190). */
int anuga_get_centroid_subset(anuga_domain *d, int64_t n, const int64_t *indices, double *values);
int anuga_set_centroid_subset(anuga_domain *d, int64_t n, const int64_t *indices,
                              const double *values);
/* Cumulative inflow volume sum(local rate * area) of operator id (order of the sum not fixed). */
int anuga_get_operator_influx(anuga_domain *d, int64_t id, double *cumulative);

/* ---- Output at yield points ---------------------------------------------------
 * Node connectivity of the mesh for SWW output: number_of_triangles_per_node
 * (nodes) and vertex_value_indices (3N, flat vertex index 3k+i, grouped by
 * node), the arrays of General_mesh.build_inverted_triangle_structure
 * (general_mesh.py:790-855). The entry order within a node is kept: it is the
 * summation order of the node average. */
int anuga_set_output_nodes(anuga_domain *d, int64_t number_of_nodes,
                           const int64_t *number_of_triangles_per_node,
                           const int64_t *vertex_value_indices);
/* One SWW frame of the current centroid state, computed on the device, as
 * float32 (the file's precision):
 * node_values [3][nodes]: stage, xmomentum, ymomentum averaged at the nodes
 *   from the centroids (Quantity.get_vertex_values with centroid averaging,
 *   quantity.py:2204-2221, quantity.c:872-914), with stage = bed and momenta
 *   0 where the node depth is below minimum_storable_height
 *   (SWW_file.store_timestep, anuga/file/sww.py:318-356);
 * centroid_values [3][N]: stage_c, xmomentum_c, ymomentum_c in the caller's
 *   order (store_quantities_centroid, sww.py:1037-1091).
 * Either pointer may be NULL. */
int anuga_get_sww_frame(anuga_domain *d, double minimum_storable_height, float *node_values,
                        float *centroid_values);
/* The static quantities of the file (quantity ELEVATION or FRICTION): node
 * averages (nodes) and centroid values (N), float32; either may be NULL
 * (SWW_file.store_connectivity, sww.py:200-258). */
int anuga_get_static_output(anuga_domain *d, int quantity, float *node_values,
                            float *centroid_values);

/* ---- Gauges: centroid time series recorded on the device ------------------------
 * Record the centroid state of `ngauges` triangles (the caller's numbering)
 * after every device-controlled step into a device buffer of `capacity` rows,
 * so that sampling every step costs no host round trip; download with
 * anuga_get_gauge_samples. The reference samples gauges from the stored SWW
 * frames (sww2csv_gauges with output_centroids, anuga/abstract_2d_finite_volumes/
 * gauge.py:84-266: the centroid values of the triangle holding each gauge);
 * this records them at every step. anuga_evolve_steps stops early when the
 * rows run out. ngauges 0 removes the gauges. */
int anuga_set_gauges(anuga_domain *d, int64_t ngauges, const int64_t *triangles,
                     int64_t capacity);
/* Record one row now, at relative_time (host-sequenced steps). */
int anuga_sample_gauges(anuga_domain *d, double relative_time);
/* Download and clear the recorded rows: *nrows rows (at most max_rows) of
 * times (relative time after the step) and values [row][gauge][w, uh, vh, z]. */
int anuga_get_gauge_samples(anuga_domain *d, int64_t max_rows, int64_t *nrows, double *times,
                            double *values);

/* protect + second-order edge extrapolation, with the result written back to
 * the centroid state. Replaces protect_new + extrapolate_second_order_edge_sw
 * (sw_domain_orig_ext.pyx:449-475; sw_domain_orig.c:1215-1281, :1994-2680).
 * mass_error (may be NULL) receives protect's added mass. */
int anuga_distribute_to_vertices_and_edges(anuga_domain *d, double *mass_error);

/* First substep of a step: protect + extrapolate + boundary + edge fluxes.
 * Replaces compute_fluxes_ext_central on substep 0 (sw_domain_orig_ext.pyx:419-432;
 * sw_domain_orig.c:765-1150). flux_timestep receives the CFL timestep
 * (1e100 when no edge carries a wave speed above epsilon). */
int anuga_compute_fluxes(anuga_domain *d, double *flux_timestep);

/* Manning friction + conserved update + fix negative cells with timestep dt:
 * replaces manning_friction_flat, Quantity.update x3 and fix_negative_cells
 * (friction.py:22-76, quantity_ext.pyx update, sw_domain_orig_ext.pyx:538-553).
 * For rk2 the result goes to the substep buffer and the step's initial state is
 * kept as the RK2 backup (backup_centroid_values, quantity.c:735-749).
 * num_negative (may be NULL) receives the fix_negative_cells count. */
int anuga_update_conserved_quantities(anuga_domain *d, double dt, int64_t *num_negative);

/* Second rk2 substep with the same dt: protect + extrapolate + boundary + fluxes
 * + friction + update + fix negative + saxpy(0.5, 0.5) with the backup
 * (generic_domain.py:2018-2045; quantity.c:752-769). */
int anuga_rk2_second_substep(anuga_domain *d, double dt, int64_t *num_negative);

/* The second and third rk3 substeps with the same dt (DE2): protect +
 * extrapolate + fluxes + friction + update, saxpy(0.25, 0.75) with the step's
 * initial state, again, then saxpy(2.0, 1.0) and the division by 3
 * (generic_domain.py:2053-2179; quantity.c:752-769). */
int anuga_rk3_later_substeps(anuga_domain *d, double dt, int64_t *num_negative);

/* ---- Per-operation calls, in the reference's own call order ------------------
 * One entry point per call that evolve_one_euler_step / evolve_one_rk2_step /
 * evolve_one_rk3_step make (generic_domain.py:1914-2179), for a binding that
 * keeps the reference's orchestration and replaces only the C extension calls:
 *   backup_conserved_quantities; distribute_to_vertices_and_edges (above);
 *   update_boundary; compute_fluxes; compute_forcing_terms; (update_timestep
 *   on the host); update_conserved_quantities = anuga_update_centroids +
 *   anuga_fix_negative_cells; saxpy_conserved_quantities; rk3's division.
 * The state stays in HBM. They are refused (ANUGA_E_STATE) between
 * anuga_update_conserved_quantities and the fused later substeps. */

/* backup_conserved_quantities: copies stage, xmomentum, ymomentum
 * (generic_domain.py:2187-2191 -> _backup_centroid_values, quantity.c:735-749). */
int anuga_backup_conserved_quantities(anuga_domain *d);

/* saxpy_conserved_quantities(a, b): q = a q + b backup for the three conserved
 * quantities (generic_domain.py:2193-2199 -> _saxpy_centroid_values,
 * quantity.c:752-769). ANUGA_E_STATE without an earlier backup. */
int anuga_saxpy_conserved_quantities(anuga_domain *d, double a, double b);

/* Q.centroid_values / divisor for the three conserved quantities: rk3's
 * closing division (generic_domain.py:2167-2170). */
int anuga_divide_conserved_quantities(anuga_domain *d, double divisor);

/* update_boundary (generic_domain.py:2288-2306): boundary states are formed
 * inside the flux kernel from the edge values, so this only checks that the
 * handle can step. Time-dependent boundary values are uploaded to slot 0 with
 * anuga_set_boundary_values before each anuga_compute_fluxes_ext_central. */
int anuga_update_boundary(anuga_domain *d);

/* compute_fluxes_ext_central(domain, timestep) (sw_domain_orig_ext.pyx:419-432
 * -> _compute_fluxes_central, sw_domain_orig.c:765-1150) over the edge values
 * of the last distribute_to_vertices_and_edges: explicit updates, the
 * substep's boundary_flux_sum, and flux_timestep = the CFL timestep on the
 * first substep of a step, else `timestep` unchanged. The substep is the count
 * of calls modulo the method's substeps, as in the reference (restarted when
 * the method changes). ANUGA_E_STATE when the edge values are stale. */
int anuga_compute_fluxes_ext_central(anuga_domain *d, double timestep, double *flux_timestep);

/* compute_forcing_terms: Manning friction (flat, or sloped with
 * use_sloped_mannings) added to the momenta's semi-implicit updates
 * (friction.py:22-76 -> _manning_friction_flat / _sloped,
 * sw_domain_orig.c:2805-2888). */
int anuga_compute_forcing_terms(anuga_domain *d);

/* Stage.update(dt), Xmom.update(dt), Ymom.update(dt)
 * (shallow_water_domain.py:2114-2118 -> _update, quantity.c:772-822): explicit
 * and semi-implicit updates; the semi-implicit updates are reset.
 * ANUGA_E_SEMI_IMPLICIT for a denominator <= 0. */
int anuga_update_centroids(anuga_domain *d, double dt);

/* fix_negative_cells (sw_domain_orig_ext.pyx:538-553 -> _orig_fix_negative_cells,
 * sw_domain_orig.c:2890-2910); num_negative (may be NULL) receives the count. */
int anuga_fix_negative_cells(anuga_domain *d, int64_t *num_negative);

/* boundary_flux_integral_operator.__call__ after a step of `timestep`
 * (anuga/operators/boundary_flux_integral_operator.py:39-49), from the
 * boundary_flux_sum of the step's substeps; read with
 * anuga_get_boundary_flux_integral. */
int anuga_boundary_flux_integral_operator(anuga_domain *d, double timestep);

/* Local extrapolation and flux updating, euler steps only
 * (set_local_extrapolation_and_flux_updating sets max_flux_update_frequency =
 * 2**nlevels, shallow_water_domain.py:1193-1215). With it above 1,
 * anuga_compute_fluxes and anuga_compute_fluxes_ext_central evaluate only the
 * edges due this step, sum the stored terms of the others, and keep the CFL
 * timestep from growing except every max_flux_update_frequency steps
 * (_compute_fluxes_central, sw_domain_orig.c:779-829). A power of two up to
 * 2^30; 1 switches it off. Setting it restarts the edge state and counters.
 * FP64, without riverwalls, low_froude or sloped friction; while it is on,
 * anuga_evolve_steps is refused (host-sequenced steps only) and so are rk2 /
 * rk3 (ANUGA_E_STATE). */
int anuga_set_max_flux_update_frequency(anuga_domain *d, int64_t max_flux_update_frequency);

/* compute_flux_update_frequency(domain, timestep) (shallow_water_domain.py:2718-2745
 * -> _compute_flux_update_frequency, sw_domain_orig.c:509-669), called after
 * update_timestep with the step's timestep (generic_domain.py:1960-1963):
 * each edge's update frequency from its edge timesteps, and the edges that
 * the next flux call evaluates. The reference keeps the step count in a
 * process-wide static; here the handle holds it. A no-op when
 * max_flux_update_frequency is 1. */
int anuga_compute_flux_update_frequency(anuga_domain *d, double timestep);

/* Timestep control of Generic_Domain.update_timestep (generic_domain.py:2349-2415). */
typedef struct anuga_time_control {
    double CFL;
    double evolve_max_timestep;
    double evolve_min_timestep;
    double relative_finaltime;   /* HUGE_VAL when there is none */
    double relative_yieldtime;   /* next yield time; HUGE_VAL when there is none */
    double fixed_flux_timestep;  /* > 0 replaces the CFL timestep (set_fixed_flux_timestep) */
    double epsilon;              /* finaltime tolerance (anuga/config.py epsilon) */
    int64_t max_smallsteps;
    int64_t default_order;
} anuga_time_control;

/* Timestep-control state carried from one run of steps to the next. */
typedef struct anuga_step_state {
    double relative_time;          /* in/out */
    double timestep;               /* out: dt of the last step */
    double flux_timestep;          /* out: CFL timestep of the last step */
    double recorded_min_timestep;  /* in/out */
    double recorded_max_timestep;  /* in/out */
    int64_t smallsteps;            /* in/out */
    int64_t order;                 /* in/out: 2, or 1 after the small-step fallback */
    int64_t steps_taken;           /* out */
    int64_t reached;               /* out: 1 final time, 2 yield time, 0 neither */
} anuga_step_state;

/* Up to max_steps full steps (evolve_one_euler_step / evolve_one_rk2_step,
 * generic_domain.py:1914-2051) with update_timestep evaluated on the device,
 * so the run needs one host synchronisation at its end. Stops early, exactly
 * where _evolve_base would yield (relative_yieldtime) or finish
 * (relative_finaltime - epsilon); the remaining launches are no-ops. One step
 * is captured once into a HIP graph and replayed (unless ANUGA_HIP_GRAPH=0 or
 * kernel timing is on). dt_history (may be NULL) receives the timesteps taken. */
int anuga_evolve_steps(anuga_domain *d, int64_t max_steps, const anuga_time_control *tc,
                       anuga_step_state *state, double *dt_history);

/* boundary_flux_sum of the last step's substeps (sums[0..2]): the boundary
 * inflow rate (sw_domain_orig.c:1119-1126). Synchronizes. */
int anuga_get_boundary_flux_sums(anuga_domain *d, double *sums);

/* Time integral of the boundary inflow accumulated on the device after every
 * update, as boundary_flux_integral_operator.__call__ does per step
 * (anuga/operators/boundary_flux_integral_operator.py:39-49). Synchronizes. */
int anuga_get_boundary_flux_integral(anuga_domain *d, double *integral);

/* Total fix_negative_cells count since creation, as of the last synchronizing
 * call (anuga_compute_fluxes synchronizes every step). */
int anuga_get_negative_cells(anuga_domain *d, int64_t *total);

/* sum(area * height_centroid) with height from the last extrapolation
 * (get_water_volume, shallow_water_domain.py:1626-1653), fixed-order FP64 reduction. */
int anuga_get_water_volume(anuga_domain *d, double *volume);

/* No counterpart in the reference: a variant of anuga_get_water_volume for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc). sum(area * max(stage -
 * elevation, 0)) over N triangles of caller-provided device arrays, reduced
 * in anuga_get_water_volume's fixed order (the same bits for the same arrays
 * in the same order), written to the device double *volume, enqueued on
 * `stream` (a hipStream_t; NULL: the null stream) without waiting. Allocates
 * nothing: its partial sums live in a static device array of the library,
 * one per GPU (4 MiB + 1 KiB, held until the process ends), so N is at most
 * 2^27 (ANUGA_E_ARG beyond) and calls that overlap on one GPU race on it.
 * Errors: anuga_last_error(NULL). */
int anuga_water_volume_noalloc(int64_t N, const double *stage, const double *elevation,
                               const double *area, double *volume, void *stream);

/* The Reverse-Cuthill-McKee triangle order applied at creation when
 * params->reorder is set: order[new] = old. Host-only (needs no device).
 * neighbours is (N,3) with boundary edges negative. */
int anuga_rcm_order(int64_t N, const int64_t *neighbours, int64_t *order);

/* The compile-time configuration of this build, chosen by the autotuner
 * (anuga_core_amd/autotune.py, recorded in anuga_core_amd/tuned.json):
 * threads per workgroup, which is also the tile size of the fused kernels,
 * and the waves per SIMD the mixed fused kernel's registers aim at.
 * Host-only; no counterpart in the reference. */
int anuga_build_config(int64_t *block, int64_t *mixed_waves);

/* The tile order applied at creation when params->reorder is 2: recursive
 * coordinate bisection of the centroids into runs of the workgroup size. Each
 * cut is across x or y and leaves a whole number of runs on its first side,
 * 40% to 60% of them and within 64 runs of the middle; with neighbours (N,3; negative on boundary edges) the
 * cut separating the fewest adjacent pairs is taken (then the one nearest the
 * middle, then x); without them, and for parts of more than 4096 runs, the
 * longer side is cut at the middle run.
 * Ties by index; each run in increasing triangle index; order[new] = old.
 * centroids is (N,2). Host-only. No counterpart in the reference (its
 * triangles are processed in the mesh's order); it exists so that runs of
 * 256 consecutive triangles are compact tiles, with short halos, for the
 * fused extrapolation + flux kernel. */
int anuga_tile_order(int64_t N, const double *centroids, const int64_t *neighbours,
                     int64_t *order);

/* The fused kernel's tile metadata, for triangles already in internal order:
 * nb is (N,3), each interior edge packed as (n << 2) | e (n the neighbour, e
 * its edge), boundary edges negative. Tile b holds triangles
 * [256 b, 256 (b + 1)). Writes start (tiles + 1 offsets into halo), halo (the
 * neighbour triangles outside each tile, each once, in order of first
 * appearance over the tile's triangles and edges; at most halo_cap), slot (N:
 * edge i's halo slot in bits 10i..10i+9, 1023 when the neighbour is in the
 * tile), *nhalo and *hcap (the largest halo of a tile; above 320 the library
 * runs the separate kernels instead).
 * Host-only; no counterpart in the reference. */
int anuga_tile_halos(int64_t N, const int32_t *nb, int64_t *start, int64_t halo_cap,
                     int32_t *halo, uint32_t *slot, int64_t *nhalo, int64_t *hcap);

/* The edge lists of the fused kernel that evaluates each edge's flux once,
 * for the same tiles and packed neighbours as anuga_tile_halos: per tile
 * (estart: tiles + 1 offsets), each edge between two tile triangles from its
 * lower-numbered side, then (from esplit[tile]) each edge to the halo from
 * the tile's side, each group ordered by edge index then triangle; an entry
 * packs the triangle's offset in the tile (bits 0-9), its edge (10-11), a
 * halo flag (12), the neighbour's tile offset or halo slot (13-22) and the
 * neighbour's edge (23-24). At most edge_cap entries; *nedges receives the
 * count. Host-only; no counterpart in the reference. */
int anuga_tile_edges(int64_t N, const int32_t *nb, int64_t *estart, int64_t *esplit,
                     int64_t edge_cap, uint32_t *edges, int64_t *nedges);

/* Wait for all work on the handle's stream. */
int anuga_synchronize(anuga_domain *d);

/* Device-side timing of the last anuga_compute_fluxes / anuga_rk2_second_substep
 * flux kernels: average milliseconds per launch of each kernel since the last
 * reset (event-timed on the handle's stream). names: "extrapolate", "flux0",
 * "update", "flux1". */
int anuga_kernel_timing(anuga_domain *d, const char *name, double *avg_ms, int64_t *launches);
int anuga_reset_kernel_timing(anuga_domain *d, int enable);

/* ---- Ensembles: many independent scenarios stepped together ----------------
 * The reference runs one Domain per process; an ensemble (flood scenarios,
 * parameter sweeps) there is a loop over Domain objects, each evolve() a
 * separate sequence of small launches. An ensemble handle holds S scenarios
 * (each its own mesh, quantities, boundary conditions and timestep control,
 * the same algorithm parameters) in one set of device buffers, so that every
 * kernel launch and every replay of the captured step graph advances all of
 * them, each with its own dt (update_timestep per scenario). Results per
 * scenario are bit-identical to a single-scenario handle of the same inputs.
 * The single-scenario entry points above refuse an ensemble handle of S > 1
 * (ANUGA_E_STATE). */

/* Allocate S scenarios of number_of_elements[s] triangles and
 * boundary_lengths[s] boundary edges; meshes follow with anuga_ensemble_set_mesh. */
int anuga_ensemble_create(int64_t nscenarios, const int64_t *number_of_elements,
                          const int64_t *boundary_lengths, const anuga_params *params,
                          anuga_domain **out);
/* Upload scenario s's mesh (validated as in anuga_create); once per scenario. */
int anuga_ensemble_set_mesh(anuga_domain *d, int64_t scenario, const anuga_mesh *mesh);
/* Per-scenario counterparts of anuga_set/get_centroid_values and anuga_set_boundary. */
int anuga_ensemble_set_centroid_values(anuga_domain *d, int64_t scenario, int quantity,
                                       const double *values);
int anuga_ensemble_get_centroid_values(anuga_domain *d, int64_t scenario, int quantity,
                                       double *values);
int anuga_ensemble_set_boundary(anuga_domain *d, int64_t scenario, const int32_t *kind,
                                const double *dirichlet_values);
/* anuga_evolve_steps for every scenario: tc[s] and states[s] per scenario. A
 * scenario that reaches its yield or final time stops there while the others
 * go on; one that is already there at the call takes no step. */
int anuga_ensemble_evolve_steps(anuga_domain *d, int64_t max_steps,
                                const anuga_time_control *tc, anuga_step_state *states);
/* anuga_distribute_to_vertices_and_edges for every scenario; mass_errors (S
 * values, may be NULL) receives each scenario's protect mass. */
int anuga_ensemble_distribute_to_vertices_and_edges(anuga_domain *d, double *mass_errors);
/* Per-scenario water volume (S values) and boundary flux integral (S values). */
int anuga_ensemble_get_water_volumes(anuga_domain *d, double *volumes);
int anuga_ensemble_get_boundary_flux_integrals(anuga_domain *d, double *integrals);
int anuga_ensemble_size(const anuga_domain *d, int64_t *nscenarios);

/* ---- Gradients: discrete adjoint of a fixed-timestep DE1 run --------------
 * The reference has no derivative of the step. These give the exact gradient
 * of the discrete FP64 DE1 scheme (branches frozen where the forward took them)
 * for calibration and data assimilation: friction and bed per triangle, and
 * the initial state, for a loss on the stage at gauge triangles. Single
 * scenario, FP64, timestep_fluxcalls == 2. */

/* Run nsteps DE1 steps with the fixed timestep dt (set_fixed_flux_timestep
 * semantics, generic_domain.py:2332-2347, with CFL 1) from the current
 * centroid state, keeping the state every checkpoint_every steps on the
 * device; gauge_stage ((nsteps / obs_every) x ngauges) receives the stage of
 * gauge_triangles (caller's numbering, distinct) after every obs_every-th
 * step. The handle's state advances to the end of the run. */
int anuga_adjoint_record(anuga_domain *d, int64_t nsteps, double dt, int64_t checkpoint_every,
                         int64_t ngauges, const int64_t *gauge_triangles, int64_t obs_every,
                         double *gauge_stage);
/* Gradient of a loss L given dL/d(gauge stage) (the record's shape) with respect
 * to the friction (N), the bed at fixed stage (N) and the initial stage,
 * xmomentum, ymomentum ((3, N); may be NULL), for the last recorded run. The
 * run's steps are recomputed segment by segment from the checkpoints. */
int anuga_adjoint_gradient(anuga_domain *d, const double *dloss_dstage, double *grad_friction,
                           double *grad_elevation, double *grad_state0);

#ifdef __cplusplus
}
#endif

#endif /* ANUGA_HIP_H */

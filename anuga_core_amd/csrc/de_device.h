// Device-side building blocks of the DE step (gfx950, FP64 path).
//
// Arithmetic is written operation for operation in the reference's order so
// that, compiled with -ffp-contract=off, each result is bit-identical to the
// reference's FP64 C (anuga/shallow_water/sw_domain_orig.c). Line citations
// are to that file unless stated otherwise.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define DE_BLOCK 256

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X. Remap
// the hardware block id so that each XCD walks one contiguous slice of the
// (locality-ordered) triangle range: neighbour gathers then hit that XCD's L2.
__device__ __forceinline__ int xcd_block()
{
    const int b = blockIdx.x, nb = gridDim.x;
    const int xcd = b & 7, slot = b >> 3;
    const int per = nb >> 3, rem = nb & 7;
    return xcd * per + (xcd < rem ? xcd : rem) + slot;
}

struct DeMesh {
    int N;
    const int *__restrict__ nbr;       // [3][N]: (n << 2) | m, or -(j+1) on boundary edge j
    const double *__restrict__ nx;     // [3][N] outward unit normal
    const double *__restrict__ ny;
    const double *__restrict__ len;    // [3][N] edge length
    const double *__restrict__ dxv;    // [3][N] edge midpoint - centroid
    const double *__restrict__ dyv;
    const double *__restrict__ cx;     // [N] centroid
    const double *__restrict__ cy;
    const double *__restrict__ area;   // [N]
    const double *__restrict__ radius; // [N]
};

struct DeParams {
    double g, eps, mah;
    double bw, bwd, buh, buhd, bvh, bvhd;
};

struct DeEdges {        // [3][N] each, written by the extrapolation
    double *we, *he, *uhe, *vhe;
};

struct DeBoundary {
    const signed char *kind;  // [M]
    const double *val;        // [3][M] Dirichlet stage, xmom, ymom
    int M;
};

// Device-resident step control: reduction results, the timestep, the
// relative time and the update_timestep state, so that a sequence of steps
// (or a captured graph of one step) needs no host round trip.
struct DeControl {
    double flux_dt;        // CFL timestep of the last first substep
    double bflux0, bflux1; // boundary inflow of the two substeps (boundary_flux_sum)
    double mass;           // protect's added mass (last reduction)
    double volume;         // last water volume reduction
    double dt;             // timestep of the current step
    double t;              // relative time
    double integral;       // boundary flux integral
    double rec_min, rec_max;
    // update_timestep parameters (generic_domain.py:2349-2415)
    double cfl, max_dt, min_dt, finaltime, yieldtime, fixed_flux_dt, eps;
    long long nsteps;      // steps completed since the last set-up
    int smallsteps, order, default_order, max_smallsteps;
    int done;              // 1: finaltime reached, 2: yield time reached (steps become no-ops)
    int status;            // error bits, DE_ERR_*
    int hist_cap, rk2;
};

#define DE_ERR_NO_NEIGHBOUR 1
#define DE_ERR_SEMI_IMPLICIT 2
#define DE_ERR_SMALL_TIMESTEP 4
#define DE_ERR_OVERSHOOT 8

// flags written by the extrapolation, per triangle
#define DE_F_CLAMP 1u   // protect raised the stage to the bed (sw_domain_orig.c:1262-1274)
#define DE_F_KEEP 2u    // centroid momentum survives protect + extrapolation

// protect (sw_domain_orig.c:1245-1276) decides the stage clamp
__device__ __forceinline__ bool de_clamped(double w, double z, double mah)
{
    const double hc = w - z;
    return (hc < mah * 1.0) && (hc <= 0.0) && (w < z);
}

// height_centroid_values after protect (sw_domain_orig.c:2020). The clamp only
// changes w when w < z, where max(w - z, 0) is 0 either way.
__device__ __forceinline__ double de_height(double w, double z)
{
    return fmax(w - z, 0.0);
}

// "surrounded by dry cells" test (sw_domain_orig.c:2046-2064) of triangle j,
// whose surrogate neighbour on a boundary edge is j itself.
__device__ __forceinline__ bool de_dry_surrounded(const DeMesh &m, const double *__restrict__ w,
                                                  const double *__restrict__ z, int j, double mah)
{
    bool dry = true;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int v = m.nbr[i * m.N + j];
        if (v >= 0) {
            const int s = v >> 2;
            dry = dry && (de_height(w[s], z[s]) < mah);
        }
    }
    return dry;
}

// Centroid stage / height / velocity of triangle j as seen by the
// extrapolation after protect and the velocity conversion
// (sw_domain_orig.c:2014-2064).
struct DeCell {
    double w, h, u, v;
};

__device__ __forceinline__ DeCell de_prepare(const DeMesh &m, const double *__restrict__ w,
                                             const double *__restrict__ uh,
                                             const double *__restrict__ vh,
                                             const double *__restrict__ z, int j, double mah,
                                             bool &keep, bool &clamp)
{
    DeCell c;
    const double wr = w[j], zj = z[j];
    clamp = de_clamped(wr, zj, mah);
    c.w = clamp ? zj : wr;
    c.h = de_height(c.w, zj);
    keep = false;
    c.u = 0.0;
    c.v = 0.0;
    if (c.h > mah) {
        keep = !de_dry_surrounded(m, w, z, j, mah);
        if (keep) {
            const double inv = 1.0 / c.h;
            c.u = uh[j] * inv;
            c.v = vh[j] * inv;
        }
    }
    return c;
}

// limit_gradient (sw_domain_orig.c:1306-1333) applied to provisional jumps d
__device__ __forceinline__ void de_limit(double d[3], double qmin, double qmax, double beta)
{
    double r = 1000.0, r0 = 1.0;
    const double tiny = 1.0e-100;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (d[i] < -tiny) r0 = qmin / d[i];
        if (d[i] > tiny) r0 = qmax / d[i];
        r = fmin(r0, r);
    }
    const double phi = fmin(r * beta, 1.0);
    d[0] = d[0] * phi;
    d[1] = d[1] * phi;
    d[2] = d[2] * phi;
}

// Auxiliary-triangle gradient, min/max bounds and limiter for one quantity
// (sw_domain_orig.c:2238-2277 and its xmom/ymom/height twins)
__device__ __forceinline__ void de_aux_edges(double qk, double q0, double q1, double q2,
                                             double dx1, double dx2, double dy1, double dy2,
                                             double inv_area2, const double dxv[3],
                                             const double dyv[3], double beta, double out[3])
{
    const double dq0 = q0 - qk;
    const double dq1 = q1 - q0;
    const double dq2 = q2 - q0;
    double a = dy2 * dq1 - dy1 * dq2;
    a *= inv_area2;
    double b = dx1 * dq2 - dx2 * dq1;
    b *= inv_area2;
    double d[3];
    d[0] = a * dxv[0] + b * dyv[0];
    d[1] = a * dxv[1] + b * dyv[1];
    d[2] = a * dxv[2] + b * dyv[2];
    const double qmax = fmax(fmax(dq0, fmax(dq0 + dq1, dq0 + dq2)), 0.0);
    const double qmin = fmin(fmin(dq0, fmin(dq0 + dq1, dq0 + dq2)), 0.0);
    de_limit(d, qmin, qmax, beta);
    out[0] = qk + d[0];
    out[1] = qk + d[1];
    out[2] = qk + d[2];
}

// Gradient toward the single interior neighbour (sw_domain_orig.c:2470-2504)
__device__ __forceinline__ void de_line_edges(double qk, double q1, double dx2, double dy2,
                                              const double dxv[3], const double dyv[3],
                                              double beta, double out[3])
{
    const double dq1 = q1 - qk;
    const double a = dq1 * dx2;
    const double b = dq1 * dy2;
    double d[3];
    d[0] = a * dxv[0] + b * dyv[0];
    d[1] = a * dxv[1] + b * dyv[1];
    d[2] = a * dxv[2] + b * dyv[2];
    double qmin, qmax;
    if (dq1 >= 0.0) {
        qmin = 0.0;
        qmax = dq1;
    } else {
        qmin = dq1;
        qmax = 0.0;
    }
    de_limit(d, qmin, qmax, beta);
    out[0] = qk + d[0];
    out[1] = qk + d[1];
    out[2] = qk + d[2];
}

// Central-upwind flux across one edge, low_froude = 0
// (sw_domain_orig.c:284-505; the rotation is _rotate, :37-57).
__device__ __forceinline__ void de_flux_central(double ql0, double ql1, double ql2, double qr0,
                                                double qr1, double qr2, double h_left,
                                                double h_right, double hle, double hre,
                                                double n1, double n2, double eps, double ze,
                                                double g, double ef[3], double &max_speed,
                                                double &pressure_flux)
{
    if (h_left == 0.0 && h_right == 0.0) {
        ef[0] = ef[1] = ef[2] = 0.0;
        max_speed = 0.0;
        pressure_flux = 0.0;
        return;
    }
    // rotate into the edge frame
    const double l1 = n1 * ql1 + n2 * ql2;
    const double l2 = -n2 * ql1 + n1 * ql2;
    const double r1 = n1 * qr1 + n2 * qr2;
    const double r2 = -n2 * qr1 + n1 * qr2;

    double ul, uhl, vhl, ur, uhr, vhr;
    uhl = l1;
    vhl = l2;
    if (hle > 0.0) {
        const double t = 1.0 / hle;
        ul = uhl * t;
        uhl = h_left * ul;
        vhl = h_left * t * vhl;
    } else {
        ul = uhl = vhl = 0.0;
    }
    uhr = r1;
    vhr = r2;
    if (hre > 0.0) {
        const double t = 1.0 / hre;
        ur = uhr * t;
        uhr = h_right * ur;
        vhr = h_right * t * vhr;
    } else {
        ur = uhr = vhr = 0.0;
    }
    const double cl = sqrt(g * h_left);
    const double cr = sqrt(g * h_right);
    double s_max = fmax(ul + cl, ur + cr);
    if (s_max < 0.0) s_max = 0.0;
    double s_min = fmin(ul - cl, ur - cr);
    if (s_min > 0.0) s_min = 0.0;

    const double fl0 = ul * h_left, fl1 = ul * uhl, fl2 = ul * vhl;
    const double fr0 = ur * h_right, fr1 = ur * uhr, fr2 = ur * vhr;
    const double denom = s_max - s_min;
    if (denom < eps) {
        ef[0] = ef[1] = ef[2] = 0.0;
        max_speed = 0.0;
        pressure_flux = 0.5 * g * 0.5 * (h_left * h_left + h_right * h_right);
        return;
    }
    max_speed = fmax(s_max, -s_min);
    const double inv = 1.0 / fmax(denom, 1.0e-100);
    const double sms = s_max * s_min;
    double e0 = s_max * fl0 - s_min * fr0;
    e0 += sms * (fmax(qr0, ze) - fmax(ql0, ze));
    e0 *= inv;
    double e1 = s_max * fl1 - s_min * fr1;
    e1 += 1.0 * sms * (uhr - uhl);
    e1 *= inv;
    double e2 = s_max * fl2 - s_min * fr2;
    e2 += 1.0 * sms * (vhr - vhl);
    e2 *= inv;
    pressure_flux = 0.5 * g * (s_max * h_left * h_left - s_min * h_right * h_right) * inv;
    // rotate back: _rotate(edgeflux, n1, -n2)
    ef[0] = e0;
    ef[1] = n1 * e1 + (-n2) * e2;
    ef[2] = -(-n2) * e1 + n1 * e2;
}

// wave64 + LDS block reductions; the combine order is fixed, so results are
// reproducible run to run.
__device__ __forceinline__ double de_wave_min(double x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x = fmin(x, __shfl_down(x, o, 64));
    return x;
}

__device__ __forceinline__ double de_wave_sum(double x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    return x;
}

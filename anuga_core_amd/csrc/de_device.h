// Device-side building blocks of the DE step (gfx950, FP64 path).
//
// Arithmetic is written operation for operation in the reference's order so
// that, compiled with -ffp-contract=off, each result is bit-identical to the
// reference's FP64 C (anuga/shallow_water/sw_domain_orig.c). Line citations
// are to that file unless stated otherwise.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// threads per workgroup of the per-triangle kernels (a multiple of 64)
#ifndef DE_BLOCK
#define DE_BLOCK 256
#endif

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

// Per-triangle records, packed so that gathering a neighbour's fields touches
// one cache line. Triangles are in the library's internal (RCM) order.
struct alignas(32) DeQ {   // conserved state + bed: one record per triangle
    double w, uh, vh, z;
};

// FP64 geometry (bit-comparable path)
struct alignas(16) DeGeoX64 {  // extrapolation geometry, formed on the host as the kernel would
    double dxv[3], dyv[3];      // edge midpoint - centroid
    // the gradient's auxiliary geometry: with at most one boundary edge
    // dx1, dx2, dy1, dy2 of the surrogate centroids (c1 - c0, c2 - c0), area2
    // and 1 / area2 (sw_domain_orig.c:2130-2160); with two boundary edges the
    // line coefficients dx2, dy2 toward the interior neighbour in aux[1], aux[3]
    // (:2421-2450)
    double aux[6];
};
struct alignas(16) DeGeoF64 {  // flux: outward normals, edge lengths, radius
    double nx[3], ny[3], len[3], radius, pad[2];
};
// mixed-precision geometry: offsets relative to the centroid, so FP32 keeps
// full relative precision on a large domain
struct alignas(16) DeGeoX32 {
    float dxv[3], dyv[3];
    float rx[3], ry[3];   // surrogate-neighbour centroid - own centroid (0 on boundary edges)
};
struct alignas(16) DeGeoF32 {
    float nx[3], ny[3], len[3], radius;
    float dzc[3];   // neighbour centroid bed - own centroid bed (0 on boundary edges)
    float pad[3];
};

// mixed-precision edge record of one triangle: per edge the stage and depth
// jumps from the FP64 centroid values and the edge momenta, plus the FP64
// centroid depth, so that the flux never subtracts two FP32 numbers of the
// size of the water depth
struct alignas(64) DeEdgeRec32 {
    float4 e[3];     // dw, dh, uh, vh
    double hc;
    double pad;
};

// one riverwall half-edge (the layout of anuga/structures/riverwall.py
// resolved per half-edge): the wall crest elevation and the hydraulic
// parameters of its wall
struct alignas(16) DeRiverwall {
    double elev, qfactor, s1, s2, h1, h2;
};

template <class T>
struct alignas(16) DeEdge4 {  // extrapolated values on one edge
    T w, h, uh, vh;
};

struct DeMesh {
    int N;                            // triangles incl. the padding between scenarios
    const int *__restrict__ scen;     // [blocks] scenario of each (logical) block; null: one scenario
    const int4 *__restrict__ nb;      // (n << 2) | m per edge, or -(j+1) on boundary edge j;
                                      // .w = number of boundary edges
    const double2 *__restrict__ cen;  // FP64 centroids
    const void *__restrict__ gx;      // DeGeoX64 or DeGeoX32
    const void *__restrict__ gf;      // DeGeoF64 or DeGeoF32
    const double *__restrict__ area;  // FP64 area (both paths)
    // FP64 flux geometry as separate arrays, [3][N] (radius [N]): every field
    // is streamed once per triangle, and separate arrays keep the flux
    // kernel's register footprint (and so its occupancy) low
    const double *__restrict__ nx;
    const double *__restrict__ ny;
    const double *__restrict__ len;
    const double *__restrict__ radius;
    // riverwalls (FP64 path): per half-edge, [3][N], the index of its
    // DeRiverwall record or -1
    const int *__restrict__ rw;
    const DeRiverwall *__restrict__ rwd;
    // the caller's index of each triangle within its scenario (the
    // reference's loop order, for the low-Froude sums)
    const int *__restrict__ caller;
    // the vertex coordinates [N][3] (x, y): sloped friction, and the fused
    // FP64 kernel's geometry where it reproduces the mesh's arrays
    const double *__restrict__ vx;
};

// Compact tiles of DE_BLOCK consecutive triangles (a space-filling-curve
// order) for the fused extrapolation + flux kernel: per tile, the neighbour
// triangles outside it (its halo), and per triangle where each interior
// edge's neighbour value sits in LDS.
struct DeTiles {
    const int *__restrict__ start;      // [tiles + 1] offsets into halo
    const int *__restrict__ halo;       // the halo triangles of each tile
    const unsigned *__restrict__ slot;  // [N] 3 x 10 bits: halo slot of edge i, 1023: in the tile
    int hcap;                           // the largest halo (LDS stride)
    // the edges whose flux a tile evaluates once (null: none built): each
    // in-tile interior edge from its lower-numbered triangle, then each edge
    // to the halo from the tile's side (DE_TE_* fields)
    const int *__restrict__ estart;     // [tiles + 1] offsets into edge
    const int *__restrict__ esplit;     // [tiles] the first edge to the halo
    const unsigned *__restrict__ edge;
};

// Local flux updating (max_flux_update_frequency > 1): per half-edge state,
// [3][N] in the internal order (half-edge (k, i) at i * N + k). e0..e2: the
// stored stage and momentum fluxes into k (length included), pg: the stored
// pressure term; et: the edge timestep radius / max(speed, eps); fuf: the
// update frequency (a power of two); acf: the flux call that last evaluated
// the edge, its sign flipped by the frequency update (the reference's
// already_computed_flux); unf: evaluate on the next call; tmin [N]: each
// triangle's smallest frequency.
struct DeLts {
    double *e0, *e1, *e2, *pg, *et;
    int *fuf, *acf, *tmin;
    unsigned char *unf;
};

// one entry of a tile's edge list: the tile triangle t and its edge i, the
// neighbour's position (tile triangle or halo slot) and edge j
#define DE_TE_T(e) ((int)((e) & 1023u))
#define DE_TE_I(e) ((int)(((e) >> 10) & 3u))
#define DE_TE_HALO(e) (((e) >> 12) & 1u)
#define DE_TE_N(e) ((int)(((e) >> 13) & 1023u))
#define DE_TE_J(e) ((int)(((e) >> 23) & 3u))
#define DE_TE_MAKE(t, i, halo, n, j) \
    ((unsigned)(t) | ((unsigned)(i) << 10) | ((unsigned)(halo) << 12) | ((unsigned)(n) << 13) | \
     ((unsigned)(j) << 23))
static_assert(DE_BLOCK <= 1024, "tile edge entries hold a 10-bit triangle index");

struct DeParams {
    double g, eps, mah;
    double bw, bwd, buh, buhd, bvh, bvhd;
    int low_froude;   // 0, 1, 2 (sw_domain_orig.c:409-425)
    int sloped;       // sloped Manning friction (sw_domain_orig.c:2840-2888)
};

struct DeBoundary {
    const signed char *kind;  // [M]
    const double *val;        // [slot][3][M] boundary stage, xmom, ymom
    int M;
    long long sstride;        // 0: one slot for every substep; 3M: a slot per substep time
};

// boundary-value slot of a flux substep: its time is t (substep 0), t + dt
// (rk2's second, rk3's second) or t + dt/2 (rk3's third), the times at which
// the reference evaluates time-dependent boundaries (generic_domain.py:1974-2179)
template <int SUBSTEP>
__device__ __forceinline__ int de_bc_slot()
{
    return SUBSTEP == 0 ? 0 : (SUBSTEP == 3 ? 2 : 1);
}

// Device-resident step control: reduction results, the timestep, the
// relative time and the update_timestep state, so that a sequence of steps
// (or a captured graph of one step) needs no host round trip.
struct DeControl {
    double flux_dt;        // CFL timestep of the last first substep
    double bflux0, bflux1, bflux2; // boundary inflow of the substeps (boundary_flux_sum)
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
    int hist_cap, method;  // method: timestep_fluxcalls (1 euler, 2 rk2, 3 rk3)
    // the scenario's triangles are [kbeg, kend) and its blocks [bbeg, bend) of
    // the internal numbering (fixed at creation)
    int kbeg, kend, bbeg, bend;
    // gauge recording (one scenario): nsteps at the last recorded row, rows
    // recorded since the last download, row capacity (0: no gauges)
    long long gstep;
    int grows, gcap;
};

// control block of the scenario that owns logical block lb
__device__ __forceinline__ DeControl *de_ctl(DeControl *ctl, const int *scen, int lb)
{
    return scen ? ctl + scen[lb] : ctl;
}

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

// Centroid stage / height / velocity as seen by the extrapolation after
// protect and the velocity conversion (sw_domain_orig.c:2014-2064).
struct DeCell {
    double w, h, u, v;
};

// Division, reciprocal and square root in the extrapolation / flux type T:
// IEEE operations for double (the FP64 path reproduces the reference's bits);
// for the mixed path's float, the hardware reciprocal and square root
// (v_rcp_f32, v_sqrt_f32: 1 ulp), a quarter of the instructions of the
// correctly rounded sequences. The mixed path's error is stated against the
// FP64 path, not bitwise.
__device__ __forceinline__ double de_div(double a, double b) { return a / b; }
__device__ __forceinline__ float de_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double de_rcp(double a) { return 1.0 / a; }
__device__ __forceinline__ float de_rcp(float a) { return __builtin_amdgcn_rcpf(a); }
__device__ __forceinline__ double de_sqrt(double a) { return sqrt(a); }
__device__ __forceinline__ float de_sqrt(float a) { return __builtin_amdgcn_sqrtf(a); }

// limit_gradient (sw_domain_orig.c:1306-1333) applied to provisional jumps d.
// The two ratios exclude each other: the numerator is selected first, so a
// wave whose lanes take both signs divides once per edge, not twice.
template <class T>
__device__ __forceinline__ void de_limit(T d[3], T qmin, T qmax, T beta)
{
    T r = T(1000.0), r0 = T(1.0);
    const T tiny = T(1.0e-100);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const T num = d[i] < -tiny ? qmin : qmax;
        if (d[i] < -tiny || d[i] > tiny) r0 = de_div(num, d[i]);
        r = fmin(r0, r);
    }
    const T phi = fmin(r * beta, T(1.0));
    d[0] = d[0] * phi;
    d[1] = d[1] * phi;
    d[2] = d[2] * phi;
}

// Auxiliary-triangle gradient, min/max bounds and limiter for one quantity
// (sw_domain_orig.c:2238-2277 and its xmom/ymom/height twins). The centroid
// differences are formed in FP64; d receives the limited jumps to the edges.
template <class T>
__device__ __forceinline__ void de_aux_jumps(double qk, double q0, double q1, double q2, T dx1,
                                             T dx2, T dy1, T dy2, T inv_area2, const T dxv[3],
                                             const T dyv[3], T beta, T d[3])
{
    const T dq0 = T(q0 - qk);
    const T dq1 = T(q1 - q0);
    const T dq2 = T(q2 - q0);
    T a = dy2 * dq1 - dy1 * dq2;
    a *= inv_area2;
    T b = dx1 * dq2 - dx2 * dq1;
    b *= inv_area2;
    d[0] = a * dxv[0] + b * dyv[0];
    d[1] = a * dxv[1] + b * dyv[1];
    d[2] = a * dxv[2] + b * dyv[2];
    const T qmax = fmax(fmax(dq0, fmax(dq0 + dq1, dq0 + dq2)), T(0.0));
    const T qmin = fmin(fmin(dq0, fmin(dq0 + dq1, dq0 + dq2)), T(0.0));
    de_limit(d, qmin, qmax, beta);
}

// Gradient toward the single interior neighbour (sw_domain_orig.c:2470-2504);
// d receives the limited jumps
template <class T>
__device__ __forceinline__ void de_line_jumps(double qk, double q1, T dx2, T dy2, const T dxv[3],
                                              const T dyv[3], T beta, T d[3])
{
    const T dq1 = T(q1 - qk);
    const T a = dq1 * dx2;
    const T b = dq1 * dy2;
    d[0] = a * dxv[0] + b * dyv[0];
    d[1] = a * dxv[1] + b * dyv[1];
    d[2] = a * dxv[2] + b * dyv[2];
    T qmin, qmax;
    if (dq1 >= T(0.0)) {
        qmin = T(0.0);
        qmax = dq1;
    } else {
        qmin = dq1;
        qmax = T(0.0);
    }
    de_limit(d, qmin, qmax, beta);
}

// Central-upwind flux across one edge, low_froude = 0
// (sw_domain_orig.c:284-505; the rotation is _rotate, :37-57).
template <class T>
__device__ __forceinline__ void de_flux_central(T ql0, T ql1, T ql2, T qr0, T qr1, T qr2,
                                                T h_left, T h_right, T hle, T hre, T n1, T n2,
                                                T eps, T ze, T g, T ef[3], T &max_speed,
                                                T &pressure_flux, int low_froude = 0,
                                                bool mirrored = false)
{
    if (h_left == T(0.0) && h_right == T(0.0)) {
        ef[0] = ef[1] = ef[2] = T(0.0);
        max_speed = T(0.0);
        pressure_flux = T(0.0);
        return;
    }
    // rotate into the edge frame
    const T l1 = n1 * ql1 + n2 * ql2;
    const T l2 = -n2 * ql1 + n1 * ql2;
    const T r1 = n1 * qr1 + n2 * qr2;
    const T r2 = -n2 * qr1 + n1 * qr2;

    T ul, uhl, vhl, ur, uhr, vhr;
    uhl = l1;
    vhl = l2;
    if (hle > T(0.0)) {
        const T t = T(1.0) / hle;
        ul = uhl * t;
        uhl = h_left * ul;
        vhl = h_left * t * vhl;
    } else {
        ul = uhl = vhl = T(0.0);
    }
    uhr = r1;
    vhr = r2;
    if (hre > T(0.0)) {
        const T t = T(1.0) / hre;
        ur = uhr * t;
        uhr = h_right * ur;
        vhr = h_right * t * vhr;
    } else {
        ur = uhr = vhr = T(0.0);
    }
    const T cl = sqrt(g * h_left);
    const T cr = sqrt(g * h_right);
    // local Froude-number scaling of the momentum diffusion
    // (sw_domain_orig.c:409-425). The reference evaluates an interior edge
    // once, from its lower-numbered triangle; `mirrored` marks the other
    // side, which adds the four squares in that triangle's order.
    T local_fr = T(1.0);
    if (low_froude) {
        const T vl = hle > T(0.0) ? l2 * (T(1.0) / hle) : T(0.0);
        const T vr = hre > T(0.0) ? r2 * (T(1.0) / hre) : T(0.0);
        const T sq = mirrored ? ul * ul + ur * ur + vl * vl + vr * vr
                              : ur * ur + ul * ul + vr * vr + vl * vl;
        const T den = cl * cl + cr * cr + T(1.0e-10);
        if (low_froude == 1) {
            local_fr = sqrt(fmax(T(0.001), fmin(T(1.0), sq / den)));
        } else {
            local_fr = sqrt(sq / den);
            local_fr = sqrt(fmin(T(1.0), T(0.01) + fmax(local_fr - T(0.01), T(0.0))));
        }
    }
    T s_max = fmax(ul + cl, ur + cr);
    if (s_max < T(0.0)) s_max = T(0.0);
    T s_min = fmin(ul - cl, ur - cr);
    if (s_min > T(0.0)) s_min = T(0.0);

    const T fl0 = ul * h_left, fl1 = ul * uhl, fl2 = ul * vhl;
    const T fr0 = ur * h_right, fr1 = ur * uhr, fr2 = ur * vhr;
    const T denom = s_max - s_min;
    if (denom < eps) {
        ef[0] = ef[1] = ef[2] = T(0.0);
        max_speed = T(0.0);
        pressure_flux = T(0.5) * g * T(0.5) * (h_left * h_left + h_right * h_right);
        return;
    }
    max_speed = fmax(s_max, -s_min);
    const T inv = T(1.0) / fmax(denom, T(1.0e-100));
    const T sms = s_max * s_min;
    T e0 = s_max * fl0 - s_min * fr0;
    e0 += sms * (fmax(qr0, ze) - fmax(ql0, ze));
    e0 *= inv;
    T e1 = s_max * fl1 - s_min * fr1;
    e1 += local_fr * sms * (uhr - uhl);
    e1 *= inv;
    T e2 = s_max * fl2 - s_min * fr2;
    e2 += local_fr * sms * (vhr - vhl);
    e2 *= inv;
    pressure_flux = T(0.5) * g * (s_max * h_left * h_left - s_min * h_right * h_right) * inv;
    // rotate back: _rotate(edgeflux, n1, -n2)
    ef[0] = e0;
    ef[1] = n1 * e1 + (-n2) * e2;
    ef[2] = -(-n2) * e1 + n1 * e2;
}

// The same flux with the stage-diffusion difference supplied by the caller:
// fmax(qr0, ze) - fmax(ql0, ze) equals h_right - h_left, which the mixed path
// forms from small differences (used with T = float).
template <class T>
__device__ __forceinline__ void de_flux_central_d(T ql1, T ql2, T qr1, T qr2, T h_left, T h_right,
                                                  T hle, T hre, T n1, T n2, T eps, T diff0, T g,
                                                  T ef[3], T &max_speed, T &pressure_flux)
{
    if (h_left == T(0.0) && h_right == T(0.0)) {
        ef[0] = ef[1] = ef[2] = T(0.0);
        max_speed = T(0.0);
        pressure_flux = T(0.0);
        return;
    }
    const T l1 = n1 * ql1 + n2 * ql2;
    const T l2 = -n2 * ql1 + n1 * ql2;
    const T r1 = n1 * qr1 + n2 * qr2;
    const T r2 = -n2 * qr1 + n1 * qr2;
    T ul, uhl, vhl, ur, uhr, vhr;
    if (hle > T(0.0)) {
        const T t = de_rcp(hle);
        ul = l1 * t;
        uhl = h_left * ul;
        vhl = h_left * t * l2;
    } else {
        ul = uhl = vhl = T(0.0);
    }
    if (hre > T(0.0)) {
        const T t = de_rcp(hre);
        ur = r1 * t;
        uhr = h_right * ur;
        vhr = h_right * t * r2;
    } else {
        ur = uhr = vhr = T(0.0);
    }
    const T cl = de_sqrt(g * h_left);
    const T cr = de_sqrt(g * h_right);
    T s_max = fmax(ul + cl, ur + cr);
    if (s_max < T(0.0)) s_max = T(0.0);
    T s_min = fmin(ul - cl, ur - cr);
    if (s_min > T(0.0)) s_min = T(0.0);
    const T denom = s_max - s_min;
    if (denom < eps) {
        ef[0] = ef[1] = ef[2] = T(0.0);
        max_speed = T(0.0);
        pressure_flux = T(0.5) * g * T(0.5) * (h_left * h_left + h_right * h_right);
        return;
    }
    max_speed = fmax(s_max, -s_min);
    const T inv = de_rcp(denom);
    const T sms = s_max * s_min;
    const T e0 = (s_max * (ul * h_left) - s_min * (ur * h_right) + sms * diff0) * inv;
    const T e1 = (s_max * (ul * uhl) - s_min * (ur * uhr) + sms * (uhr - uhl)) * inv;
    const T e2 = (s_max * (ul * vhl) - s_min * (ur * vhr) + sms * (vhr - vhl)) * inv;
    pressure_flux = T(0.5) * g * (s_max * h_left * h_left - s_min * h_right * h_right) * inv;
    ef[0] = e0;
    ef[1] = n1 * e1 - n2 * e2;
    ef[2] = n2 * e1 + n1 * e2;
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

// HIP kernels of the DE step for gfx950 (FP64 and mixed FP32/FP64).
//
// One thread per triangle, triangles in a locality order (bisection tiles or
// RCM), DE_BLOCK-thread workgroups remapped XCD-contiguous. Per Euler substep:
//   de_extrapolate  protect + velocity + dry-neighbour zeroing + limited
//                   second-order edge values (sw_domain_orig.c:1215-1281,
//                   :1994-2680). Reads the raw centroid state and never
//                   writes it: the protect / zeroing outcome goes to a flag
//                   byte, so the step's initial state doubles as the RK2 backup.
//   de_flux         per-triangle gather of its three edge fluxes in local edge
//                   order (bit-identical to the reference's compute-once-and-
//                   mirror loop, sw_domain_orig.c:765-1150), boundary states
//                   evaluated in place, CFL timestep and boundary inflow as
//                   per-block partials. In the later Runge-Kutta substeps the
//                   kernel also applies friction, the update,
//                   fix_negative_cells and the substep's combination with the
//                   step's initial state, since dt is already known.
//   de_update       friction + update + fix_negative_cells
//                   (sw_domain_orig.c:2805-2838, quantity.c:772-822,
//                   sw_domain_orig.c:2890-2910).
//   de_reduce       fixed-order reduction of the per-block partials.
//   de_xflux,       extrapolation and flux fused over compact tiles (the
//   de_xflux_mixed  bisection order): a workgroup extrapolates its
//                   triangles and their halo into LDS and gathers the fluxes
//                   from there, with the same device functions as the
//                   separate kernels, so the same bits.
// No atomics touch the state, so every run is bitwise reproducible.
//
// Every kernel is a template on T, the type of the extrapolation,
// reconstruction and flux arithmetic: double for the FP64 path (bit-comparable
// with the reference), float for the mixed path. The conserved state, the
// update, the wet/dry thresholds, the CFL reduction and the volume stay FP64
// in both.
#include "de_device.h"
#include "de_pow.h"


// ---------------------------------------------------------------------------
// geometry access, per precision
// ---------------------------------------------------------------------------
template <class T> struct DeGeo;
template <> struct DeGeo<double> {
    using X = DeGeoX64;
    using F = DeGeoF64;
};
template <> struct DeGeo<float> {
    using X = DeGeoX32;
    using F = DeGeoF32;
};

// the step's centroid state after protect + momentum zeroing, from the flag
struct DeCentroid {
    double w, uh, vh, z;
};

__device__ __forceinline__ DeCentroid de_centroid(const DeQ q, unsigned f)
{
    DeCentroid c;
    c.z = q.z;
    c.w = (f & DE_F_CLAMP) ? q.z : q.w;
    c.uh = (f & DE_F_KEEP) ? q.uh : 0.0;
    c.vh = (f & DE_F_KEEP) ? q.vh : 0.0;
    return c;
}

// ---------------------------------------------------------------------------
// extrapolation
// ---------------------------------------------------------------------------
// The extrapolation of one triangle k (sw_domain_orig.c:1215-1281,
// :1994-2680): its protected centroid cell, the protect / momentum flag,
// protect's added mass and the limited jumps from the centroid to its three
// edges. Shared by the extrapolation kernel and the tiled fused kernel, so
// both produce the same bits.
template <class T>
struct DeExtrap {
    DeCell c;
    T jw[3], jh[3], ju[3], jv[3];
    bool constant, gw, gu, gv;
    unsigned flag;
    double mass;
    double z;   // the centroid bed
};

// The extrapolation geometry of triangle k from its vertices, as
// general_mesh.py forms the arrays the reference's extrapolation reads
// (:599-626 edge midpoints 0.5 (Va + Vb); sw_domain_orig.c:2112-2135 their
// offsets from the stored centroid).
// The three vertices of a triangle (its row of the per-triangle vertex
// coordinates, the same doubles as the nodes they come from).
struct DeVerts {
    double2 v0, v1, v2;
};

__device__ __forceinline__ DeVerts de_load_verts(const DeMesh &m, int k)
{
    const double2 *r = reinterpret_cast<const double2 *>(m.vx) + 3 * (size_t)k;
    return DeVerts{r[0], r[1], r[2]};
}

__device__ __forceinline__ void de_vertex_offsets(const DeMesh &m, int k, const DeVerts &w,
                                                  double dxv[3], double dyv[3])
{
    const double2 v0 = w.v0, v1 = w.v1, v2 = w.v2;
    const double2 c = m.cen[k];
    dxv[0] = 0.5 * (v1.x + v2.x) - c.x;
    dyv[0] = 0.5 * (v1.y + v2.y) - c.y;
    dxv[1] = 0.5 * (v2.x + v0.x) - c.x;
    dyv[1] = 0.5 * (v2.y + v0.y) - c.y;
    dxv[2] = 0.5 * (v0.x + v1.x) - c.x;
    dyv[2] = 0.5 * (v0.y + v1.y) - c.y;
}

// Outward normal and length of edge i of triangle k from its vertices
// (general_mesh.py:202-232: the edge opposite vertex i, (x, y) =
// V[i+2] - V[i+1], l = sqrt(x^2 + y^2), normal (y / l, -(x / l))).
__device__ __forceinline__ void de_vertex_normal(const DeMesh &m, int k, int i, double &n1,
                                                 double &n2, double &len)
{
    const double2 *r = reinterpret_cast<const double2 *>(m.vx) + 3 * (size_t)k;
    const int a = i == 2 ? 0 : i + 1, b = i == 0 ? 2 : i - 1;
    const double2 va = r[a], vb = r[b];
    const double xn = vb.x - va.x, yn = vb.y - va.y;
    len = sqrt(xn * xn + yn * yn);
    n1 = yn / len;
    n2 = -(xn / len);
}

// vk, qk, tvk: triangle k's neighbour record, state and vertices (VG),
// loaded by the caller
template <class T, bool VG = false>
__device__ __forceinline__ void de_extrap_cell(const DeMesh &m, const DeParams &p,
                                               const DeQ *__restrict__ q, int k, int4 vk,
                                               DeQ qk, const DeVerts &tvk, int *err,
                                               DeExtrap<T> &x)
{
    const double mah = p.mah;
    const int v[3] = {vk.x, vk.y, vk.z};
    const int nbound = vk.w;
    x.z = qk.z;
    int sur[3];
#pragma unroll
    for (int i = 0; i < 3; i++) sur[i] = v[i] >= 0 ? (v[i] >> 2) : k;
    // surrogate neighbours: one record each
    DeQ qs[3];
#pragma unroll
    for (int i = 0; i < 3; i++) qs[i] = q[sur[i]];
    double hs[3];
#pragma unroll
    for (int i = 0; i < 3; i++) hs[i] = de_height(qs[i].w, qs[i].z);

    // protect + velocity of triangle k (sw_domain_orig.c:1245-1276, :2014-2064)
    const bool clamp = de_clamped(qk.w, qk.z, mah);
    DeCell &c = x.c;
    c.w = clamp ? qk.z : qk.w;
    c.h = de_height(c.w, qk.z);
    bool dry_k = true;
#pragma unroll
    for (int i = 0; i < 3; i++) dry_k = dry_k && (v[i] < 0 || hs[i] < mah);
    const bool keep = (c.h > mah) && !dry_k;
    c.u = 0.0;
    c.v = 0.0;
    if (keep) {
        const double inv = 1.0 / c.h;
        c.u = qk.uh * inv;
        c.v = qk.vh * inv;
    }
    x.mass = clamp ? (qk.z - qk.w) * m.area[k] : 0.0;
    x.flag = (clamp ? DE_F_CLAMP : 0u) | (keep ? DE_F_KEEP : 0u);

    // the same for the surrogate neighbours. A wet neighbour n keeps its
    // velocity unless all of its own neighbours are dry; k is one of
    // them, so the heights of n's other neighbours are needed only when
    // k itself is dry (a boundary edge's surrogate is k: its test is dry_k)
    DeCell nb[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const bool cl = de_clamped(qs[i].w, qs[i].z, mah);
        nb[i].w = cl ? qs[i].z : qs[i].w;
        nb[i].h = hs[i];
        nb[i].u = 0.0;
        nb[i].v = 0.0;
        if (hs[i] > mah) {
            bool dry = v[i] < 0 ? dry_k : (c.h < mah);
            if (v[i] >= 0 && dry) {
                const int4 vs = m.nb[sur[i]];
                const int vv[3] = {vs.x, vs.y, vs.z};
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    if (vv[r] >= 0) {
                        const DeQ t = q[vv[r] >> 2];
                        dry = dry && (de_height(t.w, t.z) < mah);
                    }
                }
            }
            if (!dry) {
                const double inv = 1.0 / hs[i];
                nb[i].u = qs[i].uh * inv;
                nb[i].v = qs[i].vh * inv;
            }
        }
    }

    typename DeGeo<T>::X gx;
    T dxv[3], dyv[3];
    if constexpr (VG) {
        de_vertex_offsets(m, k, tvk, dxv, dyv);
    } else {
        gx = static_cast<const typename DeGeo<T>::X *>(m.gx)[k];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            dxv[i] = gx.dxv[i];
            dyv[i] = gx.dyv[i];
        }
    }
    // limited jumps from the centroid to the edges, per quantity
    T(&jw)[3] = x.jw, (&jh)[3] = x.jh, (&ju)[3] = x.ju, (&jv)[3] = x.jv;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        jw[i] = T(0.0);
        jh[i] = T(0.0);
        ju[i] = T(0.0);
        jv[i] = T(0.0);
    }
    bool &constant = x.constant, &gw = x.gw, &gu = x.gu, &gv = x.gv;
    constant = (nbound == 3);
    gw = gu = gv = true;   // false: beta == 0, edge = centroid value
    if (nbound <= 1) {
        T dx1, dx2, dy1, dy2, area2, inv_area2;
        if constexpr (VG) {
            // the surrogate centroids' triangle (sw_domain_orig.c:2130-2160)
            const double2 c0 = m.cen[sur[0]], c1 = m.cen[sur[1]], c2 = m.cen[sur[2]];
            dx1 = c1.x - c0.x;
            dx2 = c2.x - c0.x;
            dy1 = c1.y - c0.y;
            dy2 = c2.y - c0.y;
            area2 = dy2 * dx1 - dy1 * dx2;
            inv_area2 = 1.0 / area2;
        } else if constexpr (sizeof(T) == 8) {
            dx1 = gx.aux[0];
            dx2 = gx.aux[1];
            dy1 = gx.aux[2];
            dy2 = gx.aux[3];
            area2 = gx.aux[4];
            inv_area2 = gx.aux[5];
        } else {
            dx1 = gx.rx[1] - gx.rx[0];
            dx2 = gx.rx[2] - gx.rx[0];
            dy1 = gx.ry[1] - gx.ry[0];
            dy2 = gx.ry[2] - gx.ry[0];
            area2 = dy2 * dx1 - dy1 * dx2;
            inv_area2 = de_rcp(area2);
        }
        if (area2 <= T(0.0)) {
            constant = true;
        } else {
            const T a_tmp = T(0.3), b_tmp = T(0.1);
            const T c_tmp = T(1.0) / (a_tmp - b_tmp);
            const T d_tmp = T(1.0) - (c_tmp * a_tmp);
            const T tmah = T(mah);
            const T hc = T(c.h), h0 = T(nb[0].h), h1 = T(nb[1].h), h2 = T(nb[2].h);
            const T hmin = fmin(fmin(h0, fmin(h1, h2)), hc);
            const T hmax = fmax(fmax(h0, fmax(h1, h2)), hc);
            T hfactor = fmax(T(0.0), fmin(de_div(c_tmp * fmax(hmin, T(0.0)), fmax(hc, T(1.0e-06))) + d_tmp,
                                          fmin(de_div(c_tmp * fmax(hc, T(0.0)), fmax(hmax, T(1.0e-06))) + d_tmp,
                                               T(1.0))));
            hfactor = fmin(de_div(T(1.2) * fmax(hmin - tmah, T(0.0)), fmax(hmin, T(0.0)) + T(1.0) * tmah),
                           hfactor);

            T beta = T(p.bwd) + (T(p.bw) - T(p.bwd)) * hfactor;
            gw = beta > T(0.0);
            if (gw) {
                de_aux_jumps<T>(c.w, nb[0].w, nb[1].w, nb[2].w, dx1, dx2, dy1, dy2, inv_area2,
                                dxv, dyv, beta, jw);
                de_aux_jumps<T>(c.h, nb[0].h, nb[1].h, nb[2].h, dx1, dx2, dy1, dy2, inv_area2,
                                dxv, dyv, beta, jh);
            }
            beta = T(p.buhd) + (T(p.buh) - T(p.buhd)) * hfactor;
            gu = beta > T(0.0);
            if (gu)
                de_aux_jumps<T>(c.u, nb[0].u, nb[1].u, nb[2].u, dx1, dx2, dy1, dy2, inv_area2,
                                dxv, dyv, beta, ju);
            beta = T(p.bvhd) + (T(p.bvh) - T(p.bvhd)) * hfactor;
            gv = beta > T(0.0);
            if (gv)
                de_aux_jumps<T>(c.v, nb[0].v, nb[1].v, nb[2].v, dx1, dx2, dy1, dy2, inv_area2,
                                dxv, dyv, beta, jv);
        }
    } else if (nbound == 2) {
        // the first edge whose surrogate is not k itself (sw_domain_orig.c:2421-2439)
        // selects rather than indexing, so the register arrays stay in registers
        const bool e0 = v[0] >= 0, e1 = v[1] >= 0, e2 = v[2] >= 0;
        if (!(e0 || e1 || e2)) {
            atomicOr(err, DE_ERR_NO_NEIGHBOUR);
            constant = true;
        } else {
            // field-wise selects on named copies: indexing the array here
            // would place it in private (scratch) memory
            const DeCell a0 = nb[0], a1 = nb[1], a2 = nb[2];
            DeCell n1;
            n1.w = e0 ? a0.w : (e1 ? a1.w : a2.w);
            n1.h = e0 ? a0.h : (e1 ? a1.h : a2.h);
            n1.u = e0 ? a0.u : (e1 ? a1.u : a2.u);
            n1.v = e0 ? a0.v : (e1 ? a1.v : a2.v);
            T dx2, dy2;
            if constexpr (VG) {
                // toward the one interior neighbour (sw_domain_orig.c:2421-2450)
                const int s1 = e0 ? sur[0] : (e1 ? sur[1] : sur[2]);
                const double2 cs = m.cen[s1], ck = m.cen[k];
                const double ddx = cs.x - ck.x, ddy = cs.y - ck.y;
                const double a2 = ddx * ddx + ddy * ddy;
                dx2 = 1.0 / a2;
                dy2 = dx2 * ddy;
                dx2 *= ddx;
            } else if constexpr (sizeof(T) == 8) {
                dx2 = gx.aux[1];
                dy2 = gx.aux[3];
            } else {
                const T dx1 = e0 ? gx.rx[0] : (e1 ? gx.rx[1] : gx.rx[2]);
                const T dy1 = e0 ? gx.ry[0] : (e1 ? gx.ry[1] : gx.ry[2]);
                const T area2 = dx1 * dx1 + dy1 * dy1;
                dx2 = de_rcp(area2);
                dy2 = dx2 * dy1;
                dx2 *= dx1;
            }
            const T bw = T(p.bw);
            de_line_jumps<T>(c.w, n1.w, dx2, dy2, dxv, dyv, bw, jw);
            de_line_jumps<T>(c.h, n1.h, dx2, dy2, dxv, dyv, bw, jh);
            de_line_jumps<T>(c.u, n1.u, dx2, dy2, dxv, dyv, bw, ju);
            de_line_jumps<T>(c.v, n1.v, dx2, dy2, dxv, dyv, bw, jv);
        }
    }
}

// the same, loading triangle k's records
template <class T, bool VG = false>
__device__ __forceinline__ void de_extrap_cell(const DeMesh &m, const DeParams &p,
                                               const DeQ *__restrict__ q, int k, int *err,
                                               DeExtrap<T> &x)
{
    const DeVerts tvk = VG ? de_load_verts(m, k) : DeVerts{};
    de_extrap_cell<T, VG>(m, p, q, k, m.nb[k], q[k], tvk, err, x);
}

// FP64 edge values of edge i (compile-time i, or a select on a run-time one):
// edge = centroid + limited jump; the beta == 0 and constant branches copy the
// centroid value (sw_domain_orig.c:2278-2283)
__device__ __forceinline__ void de_edge64(const DeExtrap<double> &x, int i, double &ew,
                                          double &eh, double &euh, double &evh)
{
    const double jw = i == 0 ? x.jw[0] : (i == 1 ? x.jw[1] : x.jw[2]);
    const double jh = i == 0 ? x.jh[0] : (i == 1 ? x.jh[1] : x.jh[2]);
    const double ju = i == 0 ? x.ju[0] : (i == 1 ? x.ju[1] : x.ju[2]);
    const double jv = i == 0 ? x.jv[0] : (i == 1 ? x.jv[1] : x.jv[2]);
    ew = (x.constant || !x.gw) ? x.c.w : x.c.w + jw;
    eh = (x.constant || !x.gw) ? x.c.h : x.c.h + jh;
    const double eu = (x.constant || !x.gu) ? x.c.u : x.c.u + ju;
    const double ev = (x.constant || !x.gv) ? x.c.v : x.c.v + jv;
    euh = eu * eh;
    evh = ev * eh;
}

// Mixed-path edge record of edge i (compile-time i): FP32 stage and depth
// jumps and edge momenta, referred to the FP64 centroid depth x.c.h
__device__ __forceinline__ float4 de_edge32(const DeExtrap<float> &x, int i)
{
    const float jw = i == 0 ? x.jw[0] : (i == 1 ? x.jw[1] : x.jw[2]);
    const float jh = i == 0 ? x.jh[0] : (i == 1 ? x.jh[1] : x.jh[2]);
    const float ju = i == 0 ? x.ju[0] : (i == 1 ? x.ju[1] : x.ju[2]);
    const float jv = i == 0 ? x.jv[0] : (i == 1 ? x.jv[1] : x.jv[2]);
    const float hw = x.constant ? 0.f : jw;
    const float hh = x.constant ? 0.f : jh;
    const float eh = float(x.c.h + double(hh));
    const float eu = float(x.c.u + (x.constant ? 0.0 : double(ju)));
    const float ev = float(x.c.v + (x.constant ? 0.0 : double(jv)));
    return make_float4(hw, hh, eu * eh, ev * eh);
}

template <class T>
__global__ void __launch_bounds__(DE_BLOCK)
de_extrapolate(DeMesh m, DeParams p, const DeQ *__restrict__ q, void *__restrict__ E,
               unsigned char *__restrict__ flags, double *__restrict__ part_mass,
               DeControl *ctl)
{
    __shared__ double red[DE_BLOCK / 64];
    const int lb = xcd_block();
    DeControl *cb = de_ctl(ctl, m.scen, lb);
    if (cb->done) return;
    int *err = &cb->status;
    const int N = m.N;
    const int k = lb * DE_BLOCK + threadIdx.x;
    double mass = 0.0;

    if (k < cb->kend) {
        DeExtrap<T> x;
        de_extrap_cell<T>(m, p, q, k, err, x);
        mass = x.mass;
        flags[k] = x.flag;
        if constexpr (sizeof(T) == 8) {
            // FP64 edge values: four [edge][triangle] arrays (w, h, uh, vh)
            double *__restrict__ ow = reinterpret_cast<double *>(E);
            double *__restrict__ oh = ow + 3 * (size_t)N;
            double *__restrict__ ouh = ow + 6 * (size_t)N;
            double *__restrict__ ovh = ow + 9 * (size_t)N;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                double ew, eh, euh, evh;
                de_edge64(x, i, ew, eh, euh, evh);
                const int ki = i * N + k;
                ow[ki] = ew;
                oh[ki] = eh;
                ouh[ki] = euh;
                ovh[ki] = evh;
            }
        } else {
            DeEdgeRec32 rec;
#pragma unroll
            for (int i = 0; i < 3; i++) rec.e[i] = de_edge32(x, i);
            rec.hc = x.c.h;
            rec.pad = 0.0;
            reinterpret_cast<DeEdgeRec32 *>(E)[k] = rec;
        }
    }

    // block partial of protect's added mass
    mass = de_wave_sum(mass);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = mass;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = red[0];
        for (int i = 1; i < DE_BLOCK / 64; i++) t += red[i];
        part_mass[lb] = t;
    }
}

// Write the protect / momentum-zeroing outcome back into the centroid state
// (what the reference's distribute_to_vertices_and_edges leaves behind).
__global__ void __launch_bounds__(DE_BLOCK)
de_materialize(int N, DeQ *q, const unsigned char *__restrict__ flags)
{
    const int k = xcd_block() * DE_BLOCK + threadIdx.x;
    if (k >= N) return;
    const unsigned f = flags[k];
    if (f & DE_F_CLAMP) q[k].w = q[k].z;
    if (!(f & DE_F_KEEP)) {
        q[k].uh = 0.0;
        q[k].vh = 0.0;
    }
}

// ---------------------------------------------------------------------------
// friction + update + fix_negative_cells for one triangle
// ---------------------------------------------------------------------------
__device__ __forceinline__ double de_update_one(double q, double ex, double si, double dt,
                                                int *err)
{
    // quantity.c:786-812
    if (q == 0.0) si = 0.0;
    else si /= q;
    q += dt * ex;
    const double den = 1.0 - dt * si;
    if (den <= 0.0) {
        atomicOr(err, DE_ERR_SEMI_IMPLICIT);
        return q;
    }
    return q / den;
}

// bed terms of sloped friction for one triangle
struct DeSlope {
    bool on;
    double zs, zmean;
};

// The vertex beds of the last extrapolation (edge bed z_e = w_e - h_e, vertex
// = sum of the two other edges minus the opposite one, sw_domain_orig.c:2660-2668)
// and their plane's gradient (_gradient, anuga/utilities/util_ext.h:52-92).
__device__ __forceinline__ DeSlope de_sloped_bed(const double *__restrict__ E, int N, int k,
                                                 const double *__restrict__ vx)
{
    const double *Ew = E, *Eh = E + 3 * (size_t)N;
    double ze[3];
#pragma unroll
    for (int i = 0; i < 3; i++) ze[i] = Ew[i * N + k] - Eh[i * N + k];
    const double z0 = ze[1] + ze[2] - ze[0];
    const double z1 = ze[0] + ze[2] - ze[1];
    const double z2 = ze[0] + ze[1] - ze[2];
    const double *x = vx + 6 * (size_t)k;
    const double x0 = x[0], y0 = x[1], x1 = x[2], y1 = x[3], x2 = x[4], y2 = x[5];
    const double det = (y2 - y0) * (x1 - x0) - (y1 - y0) * (x2 - x0);
    double a = (y2 - y0) * (z1 - z0) - (y1 - y0) * (z2 - z0);
    a /= det;
    double b = (x1 - x0) * (z2 - z0) - (x2 - x0) * (z1 - z0);
    b /= det;
    DeSlope s;
    s.on = true;
    s.zs = sqrt(1.0 + a * a + b * b);
    s.zmean = (z0 + z1 + z2) * (1.0 / 3.0);
    return s;
}

// The semi-implicit Manning coefficient S of one triangle:
// _manning_friction_flat with eps = minimum_allowed_height (friction.py:64-72),
// or _manning_friction_sloped (sw_domain_orig.c:2840-2888): depth over the
// mean vertex bed and the slope factor sqrt(1 + |grad z|^2). False where the
// reference adds nothing (n <= eps or a depth below eps).
__device__ __forceinline__ bool de_manning_S(const DeCentroid &c, double eta, const DeParams &p,
                                             const DeSlope &sl, double &S)
{
    if (!(eta > p.mah)) return false;
    const double h = c.w - (sl.on ? sl.zmean : c.z);
    if (!(h >= p.mah)) return false;
    S = sl.on ? -p.g * eta * eta * sl.zs * sqrt((c.uh * c.uh + c.vh * c.vh))
              : -p.g * eta * eta * sqrt((c.uh * c.uh + c.vh * c.vh));
    S /= de_pow_7_3(h);
    return true;
}

__device__ __forceinline__ DeQ de_advance(const DeCentroid c, double eta, double exw, double exuh,
                                          double exvh, double dt, const DeParams &p, int *err,
                                          unsigned long long *nneg,
                                          const DeSlope sl = DeSlope{false, 0.0, 0.0})
{
    double si_uh = 0.0, si_vh = 0.0, S;
    if (de_manning_S(c, eta, p, sl, S)) {
        si_uh += S * c.uh;
        si_vh += S * c.vh;
    }
    DeQ r;
    r.w = de_update_one(c.w, exw, 0.0, dt, err);
    r.uh = de_update_one(c.uh, exuh, si_uh, dt, err);
    r.vh = de_update_one(c.vh, exvh, si_vh, dt, err);
    r.z = c.z;
    // _orig_fix_negative_cells (every triangle is a full triangle here)
    if (r.w - c.z < 0.0) {
        atomicAdd(nneg, 1ull);
        r.w = c.z;
        r.uh = 0.0;
        r.vh = 0.0;
    }
    return r;
}

// The Runge-Kutta combination of substep SUBSTEP of the advanced state r with
// the step's initial state b (saxpy_centroid_values, quantity.c:752-769, with
// the weights of generic_domain.py:2018-2045 and :2053-2179):
// 1: rk2 end, 0.5 r + 0.5 b; 2: rk3 middle, 0.25 r + 0.75 b;
// 3: rk3 end, (2 r + 1 b) / 3.
template <int SUBSTEP>
__device__ __forceinline__ DeQ de_rk_combine(const DeQ &r, const DeQ &b)
{
    DeQ o;
    if constexpr (SUBSTEP == 1) {
        o.w = 0.5 * r.w + 0.5 * b.w;
        o.uh = 0.5 * r.uh + 0.5 * b.uh;
        o.vh = 0.5 * r.vh + 0.5 * b.vh;
    } else if constexpr (SUBSTEP == 2) {
        o.w = 0.25 * r.w + 0.75 * b.w;
        o.uh = 0.25 * r.uh + 0.75 * b.uh;
        o.vh = 0.25 * r.vh + 0.75 * b.vh;
    } else {
        o.w = (2.0 * r.w + 1.0 * b.w) / 3.0;
        o.uh = (2.0 * r.uh + 1.0 * b.uh) / 3.0;
        o.vh = (2.0 * r.vh + 1.0 * b.vh) / 3.0;
    }
    o.z = b.z;
    return o;
}

// ---------------------------------------------------------------------------
// riverwalls
// ---------------------------------------------------------------------------
// Blend the edge flux (normal frame, per unit length) towards the
// Villemonte-corrected weir discharge Qfactor 2/3 H sqrt(2/3 g H) (1 - Q2/Q1)^0.385,
// returning to the shallow-water flux between submergence s1 and s2 or
// tailwater-head / weir-height h1 and h2; the momentum fluxes take the mass
// flux's scaling (at most 10). sw_domain_orig.c:671-763, operation for operation.
__device__ __forceinline__ void de_weir_adjust(double ef[3], double hl, double hr, double g,
                                               double weir_height, const DeRiverwall &rw,
                                               double &speed)
{
    const double two3 = 2.0 / 3.0;
    const double minhd = fmin(hl, hr), maxhd = fmax(hl, hr);
    const double q1 = rw.qfactor * two3 * maxhd * sqrt(two3 * g * maxhd);
    const double q2 = rw.qfactor * two3 * minhd * sqrt(two3 * g * minhd);
    const double q_ratio = q2 / fmax(q1, 1.0e-100);
    const double h_ratio = minhd / fmax(maxhd, 1.0e-100);
    const double head_ratio = minhd / fmax(weir_height, 1.0e-100);
    double q = q1 * de_pow_0385(1.0 - q_ratio);
    if (hr > hl) q *= -1.0;
    if ((h_ratio < rw.s2) & (head_ratio < rw.h2)) {
        const double w1 = fmin(fmax(h_ratio - rw.s1, 0.0) / (rw.s2 - rw.s1), 1.0);
        const double w2 = fmin(fmax(head_ratio - rw.h1, 0.0) / (rw.h2 - rw.h1), 1.0);
        const double blended = (q * (1.0 - w1) + w1 * ef[0]) * (1.0 - w2) + w2 * ef[0];
        double scale = (fabs(ef[0]) > 1.0e-100) ? blended / ef[0] : 0.0;
        scale = fmax(scale, 0.0);
        ef[0] = blended;
        ef[1] *= fmin(scale, 10.0);
        ef[2] *= fmin(scale, 10.0);
    }
    if (fabs(ef[0]) > 0.0) speed = sqrt(g * (maxhd + weir_height)) + fabs(ef[0] / (maxhd + 1.0e-12));
}

// The weir on a riverwall edge: only where the crest stands above both cells'
// beds and either side's first-order head over z_half is positive
// (sw_domain_orig.c:905-968). wn, zn: the neighbour's centroid stage and bed
// (interior edges); on a boundary edge the right head is hc + zr - z_half.
__device__ __forceinline__ void de_riverwall_weir(const DeRiverwall &rw, double g, bool interior,
                                                  double wk, double zc, double hc, double wn,
                                                  double zn, double zl, double zr,
                                                  double z_half, double ef[3], double &speed)
{
    const double weir_height = fmax(rw.elev - fmin(zl, zr), 0.0);
    if (!(rw.elev > fmax(zc, zn))) return;
    const double hl = fmax(wk - z_half, 0.0);
    const double hr = interior ? fmax(wn - z_half, 0.0) : fmax(hc + zr - z_half, 0.0);
    if (!((hl > 0.0) || (hr > 0.0))) return;
    de_weir_adjust(ef, hl, hr, g, weir_height, rw, speed);
}

// ---------------------------------------------------------------------------
// flux
// ---------------------------------------------------------------------------
// One edge's terms of triangle k's flux sums (the body of the reference's
// edge loop, sw_domain_orig.c:828-1004, for the half-edge (k, i)): own edge
// values ql0 (stage), hle (depth), ql1, ql2 (momenta); for an interior edge
// (vi >= 0) the neighbour's at the shared edge, for a boundary edge the
// boundary state evaluated here. Adds to the sums sw, sx, sy, the boundary
// inflow and the CFL bound. Shared by the flux kernel and the tiled fused
// kernel, so both produce the same bits. fnb: where the neighbour's flag is
// (GENERAL walls; NULL: flags[n]).
template <class T, int SUBSTEP, bool GENERAL>
__device__ __forceinline__ void de_edge_terms(const DeMesh &m, const DeParams &p,
                                              const DeBoundary &bc, const DeQ *q,
                                              const unsigned char *__restrict__ flags, int N,
                                              int k, int i, int vi, const DeCentroid &c, T hc,
                                              T zc, T ql0, T hle, T ql1, T ql2, T qr0, T hre,
                                              T qr1, T qr2, T ef[3], T &pg, T &speed,
                                              const unsigned char *fnb = nullptr)
{
    const T g = T(p.g), eps = T(p.eps);
    const T zl = ql0 - hle;
    const T n1 = m.nx[i * N + k], n2 = m.ny[i * N + k];
    T zr;
    if (vi >= 0) {
        zr = qr0 - hre;
    } else {
        const int j = -vi - 1;
        const int kind = bc.kind[j];
        const double *bv = bc.val + (size_t)de_bc_slot<SUBSTEP>() * bc.sstride;
        if (kind == 0) {          // reflective
            const T r1 = -ql1 * n1 - ql2 * n2;
            const T r2 = -ql1 * n2 + ql2 * n1;
            qr0 = ql0;
            qr1 = n1 * r1 - n2 * r2;
            qr2 = n2 * r1 + n1 * r2;
        } else if (kind == 1) {   // transmissive
            qr0 = ql0;
            qr1 = ql1;
            qr2 = ql2;
        } else if (kind == 2) {   // Dirichlet (constant or time-dependent)
            qr0 = T(bv[j]);
            qr1 = T(bv[bc.M + j]);
            qr2 = T(bv[2 * bc.M + j]);
        } else if (kind == 3) {   // transmissive from centroid values
            qr0 = T(c.w);
            qr1 = T(c.uh);
            qr2 = T(c.vh);
        } else if (kind == 4) {   // stage set, normal momentum kept, tangential 0
            const T nq = n1 * ql1 + n2 * ql2;
            qr0 = T(bv[j]);
            qr1 = nq * n1;
            qr2 = nq * n2;
        } else if (kind == 5) {   // transmissive stage, zero momentum
            qr0 = ql0;
            qr1 = T(0.0);
            qr2 = T(0.0);
        } else {                  // stage set, momentum transmissive
            qr0 = T(bv[j]);
            qr1 = ql1;
            qr2 = ql2;
        }
        zr = zl;
        hre = fmax(qr0 - zr, T(0.0));
    }
    T z_half = fmax(zl, zr);
    int wall = -1;
    bool mirrored = false;
    if constexpr (GENERAL) {
        // the wall is a thin bed step between the two half-edges
        // (sw_domain_orig.c:880-893)
        if (m.rw) {
            wall = m.rw[i * N + k];
            if (wall >= 0) z_half = fmax(m.rwd[wall].elev, z_half);
        }
        if (p.low_froude && vi >= 0) mirrored = m.caller[vi >> 2] < m.caller[k];
    }
    const T h_left = fmax(hle + zl - z_half, T(0.0));
    const T h_right = fmax(hre + zr - z_half, T(0.0));
    T pflux;
    if constexpr (GENERAL)
        de_flux_central<T>(ql0, ql1, ql2, qr0, qr1, qr2, h_left, h_right, hle, hre, n1,
                           n2, eps, z_half, g, ef, speed, pflux, p.low_froude, mirrored);
    else
        de_flux_central<T>(ql0, ql1, ql2, qr0, qr1, qr2, h_left, h_right, hle, hre, n1,
                           n2, eps, z_half, g, ef, speed, pflux);
    if constexpr (GENERAL) {
        if (wall >= 0) {
            double wn = 0.0, zn = c.z;
            if (vi >= 0) {
                const int n = vi >> 2;
                // the neighbour's protect flag: from the flag array, or from
                // the fused kernel's LDS (fnb), where a halo neighbour's is
                const DeCentroid cn = de_centroid(q[n], fnb ? *fnb : flags[n]);
                wn = cn.w;
                zn = cn.z;
            }
            de_riverwall_weir(m.rwd[wall], p.g, vi >= 0, c.w, c.z, hc, wn, zn, zl, zr,
                              z_half, ef, speed);
        }
    }
    const T length = m.len[i * N + k];
    ef[0] *= length;
    ef[1] *= length;
    ef[2] *= length;
    pg = length * (-g * T(0.5) * (h_left * h_left - hle * hle - (hle + hc) * (zl - zc)) + pflux);
}

// de_edge_terms added to triangle k's sums (sw_domain_orig.c:1107-1132, the
// edge's flux, then its pressure terms along the normal), the boundary inflow
// and the CFL bound
template <class T, int SUBSTEP, bool GENERAL>
__device__ __forceinline__ void de_edge_flux(const DeMesh &m, const DeParams &p,
                                             const DeBoundary &bc, const DeQ *q,
                                             const unsigned char *__restrict__ flags, int N,
                                             int k, int i, int vi, const DeCentroid &c, T hc,
                                             T zc, T ql0, T hle, T ql1, T ql2, T qr0, T hre,
                                             T qr1, T qr2, double &sw, double &sx, double &sy,
                                             double &bflux, double &dtmin,
                                             const unsigned char *fnb = nullptr)
{
    T ef[3], pg, speed;
    de_edge_terms<T, SUBSTEP, GENERAL>(m, p, bc, q, flags, N, k, i, vi, c, hc, zc, ql0, hle, ql1,
                                       ql2, qr0, hre, qr1, qr2, ef, pg, speed, fnb);
    const T n1 = m.nx[i * N + k], n2 = m.ny[i * N + k];
    sw += double(-ef[0]);
    sx += double(-ef[1]);
    sy += double(-ef[2]);
    if (vi < 0) bflux += double(-ef[0]);
    sx -= double(n1 * pg);
    sy -= double(n2 * pg);
    if (SUBSTEP == 0 && double(speed) > p.eps) {
        const double t = 1.0 / fmax(double(speed), p.eps);
        dtmin = fmin(dtmin, m.radius[k] * t);
    }
}

// SUBSTEP 0: writes explicit updates; SUBSTEP > 0: advances with the
// scenario's dt and writes de_rk_combine<SUBSTEP>(advanced, bk) to out (out
// may be bk, or q itself: each thread reads its own q and bk before writing).
// GENERAL: the instantiation for riverwalls, low_froude 1/2 and sloped
// friction (runtime switches), so that the common kernel carries none of
// their registers.
template <class T, int SUBSTEP, bool GENERAL = false>
__global__ void __launch_bounds__(DE_BLOCK)
de_flux(DeMesh m, DeParams p, const DeQ *q, const double *__restrict__ eta,
        const T *__restrict__ E, DeBoundary bc, const unsigned char *__restrict__ flags,
        double *__restrict__ exw, double *__restrict__ exuh, double *__restrict__ exvh,
        const DeQ *bk, DeQ *out, DeControl *ctl, double *__restrict__ part_dt,
        double *__restrict__ part_bf, unsigned long long *__restrict__ nneg)
{
    __shared__ double red_dt[DE_BLOCK / 64];
    __shared__ double red_bf[DE_BLOCK / 64];
    const int lb = xcd_block();
    DeControl *cb = de_ctl(ctl, m.scen, lb);
    if (cb->done) return;
    int *err = &cb->status;
    const double dt = SUBSTEP ? cb->dt : 0.0;
    const int N = m.N;
    const int k = lb * DE_BLOCK + threadIdx.x;
    double dtmin = 1.0e100, bflux = 0.0;

    if (k < cb->kend) {
        const DeCentroid c = de_centroid(q[k], flags[k]);
        const int4 vk = m.nb[k];
        const int v[3] = {vk.x, vk.y, vk.z};
        const T *__restrict__ Ew = E;
        const T *__restrict__ Eh = E + 3 * (size_t)N;
        const T *__restrict__ Euh = E + 6 * (size_t)N;
        const T *__restrict__ Evh = E + 9 * (size_t)N;
        const T hc = T(de_height(c.w, c.z));
        const T zc = T(c.z);
        double sw = 0.0, sx = 0.0, sy = 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int ki = i * N + k;
            const T ql0 = Ew[ki], hle = Eh[ki], ql1 = Euh[ki], ql2 = Evh[ki];
            T qr0 = T(0.0), qr1 = T(0.0), qr2 = T(0.0), hre = T(0.0);
            if (v[i] >= 0) {
                const int nm = (v[i] & 3) * N + (v[i] >> 2);
                qr0 = Ew[nm];
                qr1 = Euh[nm];
                qr2 = Evh[nm];
                hre = Eh[nm];
            }
            de_edge_flux<T, SUBSTEP, GENERAL>(m, p, bc, q, flags, N, k, i, v[i], c, hc, zc, ql0,
                                              hle, ql1, ql2, qr0, hre, qr1, qr2, sw, sx, sy,
                                              bflux, dtmin);
        }
        const double inv_area = 1.0 / m.area[k];
        sw *= inv_area;
        sx *= inv_area;
        sy *= inv_area;
        if (SUBSTEP == 0) {
            exw[k] = sw;
            exuh[k] = sx;
            exvh[k] = sy;
        } else {
            DeSlope sl{false, 0.0, 0.0};
            if constexpr (GENERAL)
                if (p.sloped) sl = de_sloped_bed((const double *)E, N, k, m.vx);
            const DeQ r = de_advance(c, eta[k], sw, sx, sy, dt, p, err, nneg, sl);
            out[k] = de_rk_combine<SUBSTEP>(r, bk[k]);
        }
    }

    dtmin = de_wave_min(dtmin);
    bflux = de_wave_sum(bflux);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        red_dt[wv] = dtmin;
        red_bf[wv] = bflux;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = red_dt[0], b = red_bf[0];
        for (int i = 1; i < DE_BLOCK / 64; i++) {
            a = fmin(a, red_dt[i]);
            b += red_bf[i];
        }
        part_dt[lb] = a;
        part_bf[lb] = b;
    }
}

// ---------------------------------------------------------------------------
// local flux updating (euler steps, FP64, the common options)
// ---------------------------------------------------------------------------
// compute_fluxes with local flux updating (sw_domain_orig.c:765-1150 with the
// skip test at :829): each triangle evaluates those of its half-edges whose
// update flag is set (both halves of an edge carry the same flag, and each
// half's terms are the bits the reference mirrors to it), stores their terms,
// edge timesteps and the call number, and sums the stored terms of all three
// edges, fresh or not (:1107-1139). The CFL partials cover the evaluated
// edges; the inflow partials every boundary edge's stored flux.
__global__ void __launch_bounds__(DE_BLOCK)
de_flux_lts(DeMesh m, DeParams p, const DeQ *q, const double *__restrict__ E, DeBoundary bc,
            const unsigned char *__restrict__ flags, double *__restrict__ exw,
            double *__restrict__ exuh, double *__restrict__ exvh, DeLts lt, int call,
            DeControl *ctl, double *__restrict__ part_dt, double *__restrict__ part_bf)
{
    __shared__ double red_dt[DE_BLOCK / 64];
    __shared__ double red_bf[DE_BLOCK / 64];
    const int lb = xcd_block();
    DeControl *cb = de_ctl(ctl, m.scen, lb);
    if (cb->done) return;
    const int N = m.N;
    const int k = lb * DE_BLOCK + threadIdx.x;
    double dtmin = 1.0e100, bflux = 0.0;

    if (k < cb->kend) {
        const DeCentroid c = de_centroid(q[k], flags[k]);
        const int4 vk = m.nb[k];
        const int v[3] = {vk.x, vk.y, vk.z};
        const double *__restrict__ Ew = E;
        const double *__restrict__ Eh = E + 3 * (size_t)N;
        const double *__restrict__ Euh = E + 6 * (size_t)N;
        const double *__restrict__ Evh = E + 9 * (size_t)N;
        const double hc = de_height(c.w, c.z);
        const double zc = c.z;
        double sw = 0.0, sx = 0.0, sy = 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int ki = i * N + k;
            double e0, e1, e2, pg;
            if (lt.unf[ki]) {
                const double ql0 = Ew[ki], hle = Eh[ki], ql1 = Euh[ki], ql2 = Evh[ki];
                double qr0 = 0.0, qr1 = 0.0, qr2 = 0.0, hre = 0.0;
                if (v[i] >= 0) {
                    const int nm = (v[i] & 3) * N + (v[i] >> 2);
                    qr0 = Ew[nm];
                    qr1 = Euh[nm];
                    qr2 = Evh[nm];
                    hre = Eh[nm];
                }
                double ef[3], speed;
                de_edge_terms<double, 0, false>(m, p, bc, q, flags, N, k, i, v[i], c, hc, zc,
                                                ql0, hle, ql1, ql2, qr0, hre, qr1, qr2, ef, pg,
                                                speed);
                e0 = -ef[0];
                e1 = -ef[1];
                e2 = -ef[2];
                const double t = 1.0 / fmax(speed, p.eps);
                const double et = m.radius[k] * t;
                lt.e0[ki] = e0;
                lt.e1[ki] = e1;
                lt.e2[ki] = e2;
                lt.pg[ki] = pg;
                lt.et[ki] = et;
                lt.acf[ki] = call;
                if (speed > p.eps) dtmin = fmin(dtmin, et);
            } else {
                e0 = lt.e0[ki];
                e1 = lt.e1[ki];
                e2 = lt.e2[ki];
                pg = lt.pg[ki];
            }
            const double n1 = m.nx[ki], n2 = m.ny[ki];
            sw += e0;
            sx += e1;
            sy += e2;
            if (v[i] < 0) bflux += e0;
            sx -= n1 * pg;
            sy -= n2 * pg;
        }
        const double inv_area = 1.0 / m.area[k];
        exw[k] = sw * inv_area;
        exuh[k] = sx * inv_area;
        exvh[k] = sy * inv_area;
    }

    dtmin = de_wave_min(dtmin);
    bflux = de_wave_sum(bflux);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        red_dt[wv] = dtmin;
        red_bf[wv] = bflux;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = red_dt[0], b = red_bf[0];
        for (int i = 1; i < DE_BLOCK / 64; i++) {
            a = fmin(a, red_dt[i]);
            b += red_bf[i];
        }
        part_dt[lb] = a;
        part_bf[lb] = b;
    }
}

// The update frequency derived from an edge pair's timesteps
// (sw_domain_orig.c:573-600): int(min(et / dt, M)) of each side, the smaller,
// clamped to [1, M] and rounded down to a power of two.
__device__ __forceinline__ int de_lts_frequency(double et_a, double et_b, bool pair, int M,
                                                double timestep)
{
    const double notSoFast = 1.0;
    long long f = (long long)fmin((et_a / timestep) * notSoFast, M * 1.0);
    if (pair) f = (long long)fmin((double)(long long)fmin(et_b / timestep * notSoFast, M * 1.0),
                                  (double)f);
    if (f < 1) f = 1;
    if (f > M) f = M;
    for (long long ii = 2; ii < M; ii *= 2)
        if (f > ii && f < 2 * ii) f = ii;
    return (int)f;
}

// _compute_flux_update_frequency, part 1 (sw_domain_orig.c:540-608), for
// triangle k's half-edges; cyc is the step count modulo M. The two halves of
// an edge hold the same frequency and already_computed_flux, so each half
// replays the reference's two visits of the pair on its own copy: with a
// positive call number the first visit derives the frequency and flips the
// sign, and a second visit, made only if the new frequency divides cyc,
// flips it back; with a negative one the first visit flips and the second
// flips back and derives it; 0 (never evaluated) changes nothing. A boundary
// edge has one visit: it flips the sign and derives the frequency if the sign
// was positive. Then part 2's first loop (:618-628): tmin[k], the smallest of
// the three.
__global__ void __launch_bounds__(DE_BLOCK)
de_lts_part1(int kend, int N, const int4 *__restrict__ nb, DeLts lt, long long cyc, int M,
             double timestep)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kend) return;
    const int4 vk = nb[k];
    const int v[3] = {vk.x, vk.y, vk.z};
    int fmin3 = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int ki = i * N + k;
        int f = lt.fuf[ki];
        if ((cyc & (long long)(f - 1)) == 0) {
            int a = lt.acf[ki];
            bool derive;
            if (v[i] >= 0) {
                derive = a != 0;
            } else {
                derive = a > 0;
                a = -a;
            }
            if (derive) {
                const double etb = v[i] >= 0 ? lt.et[(v[i] & 3) * N + (v[i] >> 2)] : 0.0;
                f = de_lts_frequency(lt.et[ki], etb, v[i] >= 0, M, timestep);
                if (v[i] >= 0 && a > 0 && (cyc & (long long)(f - 1)) != 0) a = -a;
                lt.fuf[ki] = f;
            }
            lt.acf[ki] = a;
        }
        fmin3 = i == 0 ? f : min(fmin3, f);
    }
    lt.tmin[k] = fmin3;
}

// part 2's second loop (:636-656): each half-edge takes the smaller of its
// two triangles' minima, and is evaluated on the next call when its frequency
// divides the next step count.
__global__ void __launch_bounds__(DE_BLOCK)
de_lts_part2(int kend, int N, const int4 *__restrict__ nb, DeLts lt, long long cyc_next)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kend) return;
    const int4 vk = nb[k];
    const int v[3] = {vk.x, vk.y, vk.z};
    const int tk = lt.tmin[k];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int ki = i * N + k;
        const int f = v[i] >= 0 ? min(tk, lt.tmin[v[i] >> 2]) : tk;
        lt.fuf[ki] = f;
        lt.unf[ki] = (cyc_next & (long long)(f - 1)) == 0 ? 1 : 0;
    }
}

// initial state: nothing stored, every edge evaluated, frequency 1, edge
// timesteps 1e100 (shallow_water_domain.py:404-426)
__global__ void __launch_bounds__(DE_BLOCK)
de_lts_init(int n3, DeLts lt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    lt.e0[i] = 0.0;
    lt.e1[i] = 0.0;
    lt.e2[i] = 0.0;
    lt.pg[i] = 0.0;
    lt.et[i] = 1.0e100;
    lt.fuf[i] = 1;
    lt.acf[i] = 0;
    lt.unf[i] = 1;
}

// Tiled fused extrapolation + flux (FP64, any SUBSTEP; out never aliases q).
// GENERAL: riverwalls, low_froude 1/2 and sloped friction, with the tile's
// and the halo's protect flags in LDS for the walls' neighbour centroids;
// with sloped friction each triangle also writes its edge stage and depth to
// Eg, which the sloped update and the later substeps' friction read. A workgroup extrapolates its tile's DE_BLOCK triangles and
// its halo (the neighbour triangles outside the tile, each once) into LDS, then each thread gathers its triangle's three edge fluxes
// from LDS. The edge values never reach HBM. The arithmetic is
// de_extrap_cell and de_edge_flux, as in de_extrapolate and de_flux, and the
// per-block partials (CFL, boundary inflow, protect mass) are formed in the
// same order, so the results are the same bits. The state read (q) is never
// written by this kernel: later substeps write out, the step's initial state
// (rk2, rk3's last) or a third buffer (rk3's middle).
template <int SUBSTEP, bool GENERAL = false>
__global__ void __launch_bounds__(DE_BLOCK)
de_xflux(DeMesh m, DeParams p, const DeQ *__restrict__ q, const double *__restrict__ eta,
         DeTiles tl, DeBoundary bc, unsigned char *__restrict__ flags, double *__restrict__ exw,
         double *__restrict__ exuh, double *__restrict__ exvh, const DeQ *bk, DeQ *out,
         DeControl *ctl, double *__restrict__ part_dt, double *__restrict__ part_bf,
         double *__restrict__ part_mass, unsigned long long *__restrict__ nneg,
         double *__restrict__ Eg)
{
    // [4 quantities][3 edges][DE_BLOCK] tile edge values, then [4][3][hcap] halo
    extern __shared__ double lds[];
    __shared__ double red_dt[DE_BLOCK / 64], red_bf[DE_BLOCK / 64], red_m[DE_BLOCK / 64];
    const int lb = xcd_block();
    DeControl *cb = de_ctl(ctl, m.scen, lb);
    if (cb->done) return;
    int *err = &cb->status;
    const double dt = SUBSTEP ? cb->dt : 0.0;
    const int N = m.N, t = threadIdx.x;
    const int k0 = lb * DE_BLOCK, k = k0 + t;
    const int kend = cb->kend;
    const int hcap = tl.hcap;
    double *const hl = lds + 12 * DE_BLOCK;
    // GENERAL: the protect flags of the tile's and the halo's triangles
    unsigned char *const lfl = reinterpret_cast<unsigned char *>(hl + 12 * hcap);
    double mass = 0.0, dtmin = 1.0e100, bflux = 0.0;
    unsigned fk = 0;

    if (k < kend) {
        DeExtrap<double> x;
        de_extrap_cell<double>(m, p, q, k, err, x);
        mass = x.mass;
        fk = x.flag;
        flags[k] = (unsigned char)fk;
        if constexpr (GENERAL) lfl[t] = (unsigned char)fk;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            double ew, eh, euh, evh;
            de_edge64(x, i, ew, eh, euh, evh);
            lds[(0 + i) * DE_BLOCK + t] = ew;
            lds[(3 + i) * DE_BLOCK + t] = eh;
            lds[(6 + i) * DE_BLOCK + t] = euh;
            lds[(9 + i) * DE_BLOCK + t] = evh;
            if constexpr (GENERAL) {
                // sloped friction reads the triangle's edge stage and depth
                // (its vertex beds) from HBM, in the update and below
                if (p.sloped) {
                    Eg[(size_t)i * N + k] = ew;
                    Eg[3 * (size_t)N + (size_t)i * N + k] = eh;
                }
            }
        }
    }
    // halo triangles: all three edges, so that the edge is chosen by the
    // reader's LDS address (a run-time choice between the registers of x
    // would place x in scratch)
    const int h0 = tl.start[lb], nh = tl.start[lb + 1] - h0;
    for (int j = t; j < nh; j += DE_BLOCK) {
        DeExtrap<double> x;
        de_extrap_cell<double>(m, p, q, tl.halo[h0 + j], err, x);
        if constexpr (GENERAL) lfl[DE_BLOCK + j] = (unsigned char)x.flag;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            double ew, eh, euh, evh;
            de_edge64(x, i, ew, eh, euh, evh);
            hl[(0 + i) * hcap + j] = ew;
            hl[(3 + i) * hcap + j] = eh;
            hl[(6 + i) * hcap + j] = euh;
            hl[(9 + i) * hcap + j] = evh;
        }
    }
    __syncthreads();

    if (k < kend) {
        const DeCentroid c = de_centroid(q[k], fk);
        const int4 vk = m.nb[k];
        const int v[3] = {vk.x, vk.y, vk.z};
        const unsigned sl = tl.slot[k];
        const double hc = de_height(c.w, c.z);
        const double zc = c.z;
        double sw = 0.0, sx = 0.0, sy = 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const double ql0 = lds[(0 + i) * DE_BLOCK + t], hle = lds[(3 + i) * DE_BLOCK + t];
            const double ql1 = lds[(6 + i) * DE_BLOCK + t], ql2 = lds[(9 + i) * DE_BLOCK + t];
            double qr0 = 0.0, qr1 = 0.0, qr2 = 0.0, hre = 0.0;
            const unsigned char *fnb = nullptr;
            if (v[i] >= 0) {
                const unsigned s = (sl >> (10 * i)) & 1023u;
                const double *src;
                int stride;
                if (s == 1023u) {
                    src = lds + (v[i] & 3) * DE_BLOCK + ((v[i] >> 2) - k0);
                    stride = 3 * DE_BLOCK;
                    if constexpr (GENERAL) fnb = lfl + ((v[i] >> 2) - k0);
                } else {
                    src = hl + (v[i] & 3) * hcap + s;
                    stride = 3 * hcap;
                    if constexpr (GENERAL) fnb = lfl + DE_BLOCK + s;
                }
                qr0 = src[0];
                hre = src[stride];
                qr1 = src[2 * stride];
                qr2 = src[3 * stride];
            }
            de_edge_flux<double, SUBSTEP, GENERAL>(m, p, bc, q, flags, N, k, i, v[i], c, hc, zc,
                                                   ql0, hle, ql1, ql2, qr0, hre, qr1, qr2, sw,
                                                   sx, sy, bflux, dtmin, fnb);
        }
        const double inv_area = 1.0 / m.area[k];
        sw *= inv_area;
        sx *= inv_area;
        sy *= inv_area;
        if (SUBSTEP == 0) {
            exw[k] = sw;
            exuh[k] = sx;
            exvh[k] = sy;
        } else {
            DeSlope sl{false, 0.0, 0.0};
            if constexpr (GENERAL)
                if (p.sloped) sl = de_sloped_bed(Eg, N, k, m.vx);
            const DeQ r = de_advance(c, eta[k], sw, sx, sy, dt, p, err, nneg, sl);
            out[k] = de_rk_combine<SUBSTEP>(r, bk[k]);
        }
    }

    dtmin = de_wave_min(dtmin);
    bflux = de_wave_sum(bflux);
    mass = de_wave_sum(mass);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        red_dt[wv] = dtmin;
        red_bf[wv] = bflux;
        red_m[wv] = mass;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = red_dt[0], b = red_bf[0], cm = red_m[0];
        for (int i = 1; i < DE_BLOCK / 64; i++) {
            a = fmin(a, red_dt[i]);
            b += red_bf[i];
            cm += red_m[i];
        }
        part_dt[lb] = a;
        part_bf[lb] = b;
        part_mass[lb] = cm;
    }
}

// The workgroup's partial results: its smallest CFL timestep, boundary flux
// and protect mass (substep 0), or its boundary flux alone (the later
// substeps' reductions read nothing else). A sum whose lanes are all zero is
// zero without the exchange.
__device__ __forceinline__ double de_wave_sum_nz(double x)
{
    return __ballot(x != 0.0) ? de_wave_sum(x) : 0.0;
}

template <int SUBSTEP>
__device__ __forceinline__ void de_block_partials(double dtmin, double bflux, double mass,
                                                  double *red_dt, double *red_bf, double *red_m,
                                                  double *part_dt, double *part_bf,
                                                  double *part_mass, int lb)
{
    if (SUBSTEP == 0) dtmin = de_wave_min(dtmin);
    bflux = de_wave_sum_nz(bflux);
    if (SUBSTEP == 0) mass = de_wave_sum_nz(mass);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        red_dt[wv] = dtmin;
        red_bf[wv] = bflux;
        red_m[wv] = mass;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = red_dt[0], b = red_bf[0], cm = red_m[0];
        for (int i = 1; i < DE_BLOCK / 64; i++) {
            a = fmin(a, red_dt[i]);
            b += red_bf[i];
            cm += red_m[i];
        }
        if (SUBSTEP == 0) {
            part_dt[lb] = a;
            part_mass[lb] = cm;
        }
        part_bf[lb] = b;
    }
}

// One listed edge of a tile (de_xflux_e): its flux from the tile triangle's
// side (the entry's halo bit: the neighbour is in the halo; otherwise in the
// tile, and its side's pressure term and CFL bound are formed here too),
// written over the LDS edge values that only this entry reads. The halo bit
// is read at run time, so a wave can take edges of both kinds in one pass.
template <int SUBSTEP, bool VG>
__device__ __forceinline__ void de_tile_edge(const DeMesh &m, const DeParams &p, unsigned en,
                                             int k0, double *lds, double *hl,
                                             const double *hz, int hcap, double &dtmin)
{
    constexpr int Q = 3 * DE_BLOCK;
    const double g = p.g, eps = p.eps;
    const int N = m.N;
    const int ta = DE_TE_T(en), ia = DE_TE_I(en), nb = DE_TE_N(en), jb = DE_TE_J(en);
    const bool halo = DE_TE_HALO(en);
    const int ka = k0 + ta;
    double *const A = lds + ia * DE_BLOCK + ta;
    double *const R = halo ? hl + jb * hcap + nb : lds + jb * DE_BLOCK + nb;
    const int rs = halo ? 3 * hcap : Q;
    const double ql0 = A[0], hle = A[Q], ql1 = A[2 * Q], ql2 = A[3 * Q];
    const double qr0 = R[0], hre = R[rs], qr1 = R[2 * rs], qr2 = R[3 * rs];
    double n1, n2, length;
    if constexpr (VG) {
        de_vertex_normal(m, ka, ia, n1, n2, length);
    } else {
        n1 = m.nx[ia * N + ka];
        n2 = m.ny[ia * N + ka];
        length = m.len[ia * N + ka];
    }
    const double zl = ql0 - hle, zr = qr0 - hre;
    const double z_half = fmax(zl, zr);
    const double h_left = fmax(hle + zl - z_half, 0.0);
    const double h_right = fmax(hre + zr - z_half, 0.0);
    double ef[3], speed, pflux;
    de_flux_central<double>(ql0, ql1, ql2, qr0, qr1, qr2, h_left, h_right, hle, hre, n1, n2,
                            eps, z_half, g, ef, speed, pflux);
    ef[0] *= length;
    ef[1] *= length;
    ef[2] *= length;
    const double hca = hz[ta], zca = hz[DE_BLOCK + ta];
    const double pga = length * (-g * 0.5 * (h_left * h_left - hle * hle -
                                             (hle + hca) * (zl - zca)) + pflux);
    A[0] = -ef[0];
    A[Q] = -ef[1];
    A[2 * Q] = -ef[2];
    A[3 * Q] = n1 * pga;
    R[0] = n2 * pga;
    double tt = 0.0;
    const bool cfl = SUBSTEP == 0 && speed > eps;
    if (cfl) {
        tt = 1.0 / fmax(speed, eps);
        dtmin = fmin(dtmin, m.radius[ka] * tt);
    }
    if (!halo) {
        // the other side: its own depths and bed, the same flux
        const double hcb = hz[nb], zcb = hz[DE_BLOCK + nb];
        const double pgb = length * (-g * 0.5 * (h_right * h_right - hre * hre -
                                                 (hre + hcb) * (zr - zcb)) + pflux);
        R[rs] = n1 * pgb;
        R[2 * rs] = n2 * pgb;
        if (cfl) dtmin = fmin(dtmin, m.radius[k0 + nb] * tt);
    }
}

// The fused FP64 kernel with each edge's flux evaluated once (the common
// options: no riverwalls, low_froude 0, flat friction; the mesh's normals and
// lengths mirror exactly across every interior edge). VG: the geometry is
// formed from the vertices. Phases, with a barrier after each:
//  1. every thread extrapolates its triangle into LDS (de_extrap_cell, as
//     de_xflux);
//  2. the first waves extrapolate the halo while each thread of the others
//     evaluates one edge inside the tile, from its lower-numbered side, which
//     also forms the other side's pressure term (the reference's compute-
//     once-and-mirror loop, sw_domain_orig.c:996-1009);
//  3. the tile's remaining edges and the edges to the halo (from the tile's
//     side) in one pass over all the waves;
//  4. each thread gathers its triangle's three edges in local order as
//     de_xflux does: the same operations on the same values, so the same bits
//     (the other side's flux is exactly the negation, its pressure flux and
//     wave speed identical); boundary edges are evaluated here, from the
//     untouched edge values.
// Each listed edge's results overwrite the LDS edge values only it reads.
// Four waves per SIMD (at most 128 VGPRs), which the LDS allows too.
template <int SUBSTEP, bool VG = false>
__global__ void __launch_bounds__(DE_BLOCK, 4)
de_xflux_e(DeMesh m, DeParams p, const DeQ *__restrict__ q, const double *__restrict__ eta,
           DeTiles tl, DeBoundary bc, unsigned char *__restrict__ flags, double *__restrict__ exw,
           double *__restrict__ exuh, double *__restrict__ exvh, const DeQ *bk, DeQ *out,
           DeControl *ctl, double *__restrict__ part_dt, double *__restrict__ part_bf,
           double *__restrict__ part_mass, unsigned long long *__restrict__ nneg)
{
    static_assert(DE_BLOCK >= 128, "the halo and the tile's edges share the workgroup's waves");
    // [4 quantities][3 edges][DE_BLOCK] tile edge values, [4][3][hcap] halo,
    // then [2][DE_BLOCK] the tile's centroid depths and beds
    extern __shared__ double lds[];
    __shared__ double red_dt[DE_BLOCK / 64], red_bf[DE_BLOCK / 64], red_m[DE_BLOCK / 64];
    const int lb = xcd_block();
    const int N = m.N, t = threadIdx.x;
    const int k0 = lb * DE_BLOCK, k = k0 + t;
    // phase 1's own records, requested before the control block is read
    // (bounded by the arrays' length, not yet by the scenario's end)
    int4 vk1 = make_int4(-1, -1, -1, 3);
    DeVerts tv1{};
    DeQ qk1{0.0, 0.0, 0.0, 0.0};
    if (k < N) {
        vk1 = m.nb[k];
        qk1 = q[k];
        if (VG) tv1 = de_load_verts(m, k);
    }
    // the edge-list entries of phases 2 and 3, loaded now so that their
    // latency hides behind the extrapolation: hw waves take the halo in
    // phase 2 (at least one wave left for the tile's edges), each thread of
    // the others one edge inside the tile
    const int h0 = tl.start[lb], nh = tl.start[lb + 1] - h0;
    const int e0 = tl.estart[lb], es = tl.esplit[lb], e1 = tl.estart[lb + 1];
    const int hw = min((nh + 63) >> 6, DE_BLOCK / 64 - 1);
    const int e2 = e0 + t - 64 * hw, e3 = min(es, e0 + DE_BLOCK - 64 * hw) + t;
    const bool has2 = t >= 64 * hw && e2 < es;
    // phase 2's item: a tile edge, or the first halo triangle of a halo thread
    const unsigned en2 = has2 ? tl.edge[e2] : (t < nh && t < 64 * hw ? (unsigned)tl.halo[h0 + t] : 0u);
    const unsigned en3 = e3 < e1 ? tl.edge[e3] : 0u;

    DeControl *cb = de_ctl(ctl, m.scen, lb);
    if (cb->done) return;
    int *err = &cb->status;
    const double dt = SUBSTEP ? cb->dt : 0.0;
    const int kend = cb->kend;
    const int hcap = tl.hcap;
    double *const hl = lds + 12 * DE_BLOCK;
    double *const hz = hl + 12 * hcap;
    double mass = 0.0, dtmin = 1.0e100, bflux = 0.0;
    unsigned fk = 0;
    constexpr int Q = 3 * DE_BLOCK;   // stride between quantities of the tile
    // phase 1
    if (k < kend) {
        DeExtrap<double> x;
        de_extrap_cell<double, VG>(m, p, q, k, vk1, qk1, tv1, err, x);
        mass = x.mass;
        fk = x.flag;
        flags[k] = (unsigned char)fk;
        hz[t] = x.c.h;
        hz[DE_BLOCK + t] = x.z;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            double ew, eh, euh, evh;
            de_edge64(x, i, ew, eh, euh, evh);
            lds[(0 + i) * DE_BLOCK + t] = ew;
            lds[(3 + i) * DE_BLOCK + t] = eh;
            lds[(6 + i) * DE_BLOCK + t] = euh;
            lds[(9 + i) * DE_BLOCK + t] = evh;
        }
    }
    // phase 2's records, requested before the barrier: a halo thread's
    // first triangle's neighbours
    const bool hal2 = t < nh && t < 64 * hw;
    int4 rec2 = make_int4(-1, -1, -1, 3);
    if (hal2) rec2 = m.nb[en2];
    __syncthreads();

    // phase 2: the halo's extrapolation, and edges inside the tile
    if (t < 64 * hw) {
        for (int j = t; j < nh; j += 64 * hw) {
            DeExtrap<double> x;
            const int kh = j == t ? (int)en2 : tl.halo[h0 + j];
            if (j != t) rec2 = m.nb[kh];   // a further halo triangle
            const DeVerts tvh = VG ? de_load_verts(m, kh) : DeVerts{};
            de_extrap_cell<double, VG>(m, p, q, kh, rec2, q[kh], tvh, err, x);
#pragma unroll
            for (int i = 0; i < 3; i++) {
                double ew, eh, euh, evh;
                de_edge64(x, i, ew, eh, euh, evh);
                hl[(0 + i) * hcap + j] = ew;
                hl[(3 + i) * hcap + j] = eh;
                hl[(6 + i) * hcap + j] = euh;
                hl[(9 + i) * hcap + j] = evh;
            }
        }
    } else if (has2) {
        de_tile_edge<SUBSTEP, VG>(m, p, en2, k0, lds, hl, hz, hcap, dtmin);
    }
    __syncthreads();

    // phase 3: the tile's remaining edges, then the edges to the halo
    if (e3 < e1) {
        de_tile_edge<SUBSTEP, VG>(m, p, en3, k0, lds, hl, hz, hcap, dtmin);
        for (int e = e3 + DE_BLOCK; e < e1; e += DE_BLOCK)
            de_tile_edge<SUBSTEP, VG>(m, p, tl.edge[e], k0, lds, hl, hz, hcap, dtmin);
    }
    // phase 4's inputs, requested before the barrier (held from the
    // kernel's start they cost registers in every phase)
    DeQ qk4{0.0, 0.0, 0.0, 0.0}, bk_k{0.0, 0.0, 0.0, 0.0};
    int4 vk4 = make_int4(-1, -1, -1, 3);
    double eta_k = 0.0, area_k = 1.0;
    unsigned sl_k = 0;
    if (k < kend) {
        qk4 = q[k];
        vk4 = m.nb[k];
        area_k = m.area[k];
        sl_k = tl.slot[k];
        if (SUBSTEP) {
            eta_k = eta[k];
            bk_k = bk[k];
        }
    }
    __syncthreads();

    // phase 4: each triangle's three edges in local order
    if (k < kend) {
        const DeCentroid c = de_centroid(qk4, fk);
        const int4 vk = vk4;
        const int v[3] = {vk.x, vk.y, vk.z};
        const unsigned sl = sl_k;
        double sw = 0.0, sx = 0.0, sy = 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            double *const A = lds + i * DE_BLOCK + t;
            if (v[i] < 0) {
                const double hc = de_height(c.w, c.z);
                de_edge_flux<double, SUBSTEP, false>(m, p, bc, q, flags, N, k, i, v[i], c, hc,
                                                     c.z, A[0], A[Q], A[2 * Q], A[3 * Q], 0.0,
                                                     0.0, 0.0, 0.0, sw, sx, sy, bflux, dtmin);
                continue;
            }
            const int n = v[i] >> 2, j = v[i] & 3;
            const unsigned s = (sl >> (10 * i)) & 1023u;
            double a0, a1, a2, b1, b2;
            if (s != 1023u) {           // to the halo: this side's entry
                a0 = A[0];
                a1 = A[Q];
                a2 = A[2 * Q];
                b1 = A[3 * Q];
                b2 = hl[j * hcap + s];
            } else if (k < n) {         // evaluated from this side
                a0 = A[0];
                a1 = A[Q];
                a2 = A[2 * Q];
                b1 = A[3 * Q];
                b2 = lds[j * DE_BLOCK + (n - k0)];
            } else {                    // evaluated from the neighbour's side
                const double *O = lds + j * DE_BLOCK + (n - k0);
                a0 = -O[0];
                a1 = -O[Q];
                a2 = -O[2 * Q];
                b1 = -A[Q];
                b2 = -A[2 * Q];
            }
            sw += a0;
            sx += a1;
            sy += a2;
            sx -= b1;
            sy -= b2;
        }
        const double inv_area = 1.0 / area_k;
        sw *= inv_area;
        sx *= inv_area;
        sy *= inv_area;
        if (SUBSTEP == 0) {
            exw[k] = sw;
            exuh[k] = sx;
            exvh[k] = sy;
        } else {
            const DeQ r = de_advance(c, eta_k, sw, sx, sy, dt, p, err, nneg);
            out[k] = de_rk_combine<SUBSTEP>(r, bk_k);
        }
    }

    de_block_partials<SUBSTEP>(dtmin, bflux, mass, red_dt, red_bf, red_m, part_dt, part_bf,
                               part_mass, lb);
}

// One edge's terms of triangle k's sums on the mixed path (the body of the
// reference's edge loop, sw_domain_orig.c:828-1004, over FP32 differences):
// own edge record el (dw, dh, uh, vh) and centroid depth hcd; for an interior
// edge the neighbour's record er at the shared edge and its centroid depth
// hcn. Shared by de_flux_mixed and the tiled fused kernel.
template <int SUBSTEP>
__device__ __forceinline__ void de_edge_flux_mixed(const DeParams &p, const DeBoundary &bc,
                                                   const DeGeoF32 *__restrict__ gf,
                                                   const DeCentroid &c,
                                                   double hcd, float hc, int i, int vi,
                                                   float4 el, float4 er, double hcn, double &sw,
                                                   double &sx, double &sy, double &bflux,
                                                   double &dtmin)
{
    const float g = float(p.g), eps = float(p.eps);
    const float hle = float(hcd + double(el.y));
    const float zl = el.x - el.y;
    const float ql1 = el.z, ql2 = el.w;
    const float n1 = gf->nx[i], n2 = gf->ny[i];
    float hre, zr, qr1, qr2, dH = 0.f;
    if (vi >= 0) {
        hre = float(hcn + double(er.y));
        zr = gf->dzc[i] + (er.x - er.y);
        qr1 = er.z;
        qr2 = er.w;
        dH = (float(hcn - hcd) + (er.y - el.y)) + (zr - zl);
    } else {
        const int j = -vi - 1;
        const int kind = bc.kind[j];
        const double *bv = bc.val + (size_t)de_bc_slot<SUBSTEP>() * bc.sstride;
        float wr;                                   // stage above z_c
        if (kind == 0) {          // reflective
            const float r1 = -ql1 * n1 - ql2 * n2;
            const float r2 = -ql1 * n2 + ql2 * n1;
            wr = zl + hle;
            qr1 = n1 * r1 - n2 * r2;
            qr2 = n2 * r1 + n1 * r2;
        } else if (kind == 1) {   // transmissive
            wr = zl + hle;
            qr1 = ql1;
            qr2 = ql2;
        } else if (kind == 2) {   // Dirichlet
            wr = float(bv[j] - c.z);
            qr1 = float(bv[bc.M + j]);
            qr2 = float(bv[2 * bc.M + j]);
        } else if (kind == 3) {   // transmissive from centroid values
            wr = float(c.w - c.z);
            qr1 = float(c.uh);
            qr2 = float(c.vh);
        } else if (kind == 4) {   // stage set, normal momentum kept, tangential 0
            const float nq = n1 * ql1 + n2 * ql2;
            wr = float(bv[j] - c.z);
            qr1 = nq * n1;
            qr2 = nq * n2;
        } else if (kind == 5) {   // transmissive stage, zero momentum
            wr = zl + hle;
            qr1 = 0.f;
            qr2 = 0.f;
        } else {                  // stage set, momentum transmissive
            wr = float(bv[j] - c.z);
            qr1 = ql1;
            qr2 = ql2;
        }
        zr = zl;
        hre = (kind <= 1 || kind == 5) ? fmax(hle, 0.f) : fmax(wr - zr, 0.f);
    }
    const float z_half = fmax(zl, zr);
    const float hl_raw = hle + (zl - z_half);
    const float hr_raw = hre + (zr - z_half);
    const float h_left = fmax(hl_raw, 0.f);
    const float h_right = fmax(hr_raw, 0.f);
    const float diff0 = (vi >= 0 && hl_raw > 0.f && hr_raw > 0.f) ? dH : h_right - h_left;
    float ef[3], speed, pflux;
    de_flux_central_d<float>(ql1, ql2, qr1, qr2, h_left, h_right, hle, hre, n1, n2, eps,
                             diff0, g, ef, speed, pflux);
    const float length = gf->len[i];
    const float dsq = (hl_raw > 0.f) ? (zl - z_half) * (h_left + hle) : -(hle * hle);
    const float pg = length * (-g * 0.5f * (dsq - (hle + hc) * zl) + pflux);
    sw -= double(ef[0] * length);
    sx -= double(ef[1] * length);
    sy -= double(ef[2] * length);
    if (vi < 0) bflux -= double(ef[0] * length);
    sx -= double(n1 * pg);
    sy -= double(n2 * pg);
    if (SUBSTEP == 0 && double(speed) > p.eps)
        dtmin = fmin(dtmin, double(gf->radius) * (1.0 / double(speed)));
}

// Mixed-precision flux. All FP32 arithmetic is on quantities referred to the
// own centroid bed z_c (a local datum): edge bed zl = dw - dh, neighbour edge
// bed zr = (z_n - z_c) + dw' - dh', depths h = h_c + dh formed in FP64 and
// rounded, and the two terms that are differences of depths,
//   fmax(wr, z_half) - fmax(wl, z_half) = h_right - h_left
//   h_left^2 - hle^2 = (zl - z_half)(h_left + hle)   (h_left > 0),
// evaluated from small differences. Equal to the reference's formulas in real
// arithmetic (sw_domain_orig.c:879-1004); sums, CFL and update stay FP64.
template <int SUBSTEP>
__global__ void __launch_bounds__(DE_BLOCK)
de_flux_mixed(DeMesh m, DeParams p, const DeQ *q, const double *__restrict__ eta,
              const DeEdgeRec32 *__restrict__ E, DeBoundary bc,
              const unsigned char *__restrict__ flags, double *__restrict__ exw,
              double *__restrict__ exuh, double *__restrict__ exvh, const DeQ *bk, DeQ *out,
              DeControl *ctl, double *__restrict__ part_dt, double *__restrict__ part_bf,
              unsigned long long *__restrict__ nneg)
{
    __shared__ double red_dt[DE_BLOCK / 64];
    __shared__ double red_bf[DE_BLOCK / 64];
    const int lb = xcd_block();
    DeControl *cb = de_ctl(ctl, m.scen, lb);
    if (cb->done) return;
    int *err = &cb->status;
    const double dt = SUBSTEP ? cb->dt : 0.0;
    const int N = m.N;
    const int k = lb * DE_BLOCK + threadIdx.x;
    double dtmin = 1.0e100, bflux = 0.0;

    if (k < cb->kend) {
        const DeCentroid c = de_centroid(q[k], flags[k]);
        const int4 vk = m.nb[k];
        const int v[3] = {vk.x, vk.y, vk.z};
        const DeGeoF32 *gf = static_cast<const DeGeoF32 *>(m.gf) + k;
        const DeEdgeRec32 own = E[k];
        const double hcd = own.hc;
        const float hc = float(hcd);
        double sw = 0.0, sx = 0.0, sy = 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            float4 er = make_float4(0.f, 0.f, 0.f, 0.f);
            double hcn = 0.0;
            if (v[i] >= 0) {
                const DeEdgeRec32 *rn = E + (v[i] >> 2);
                er = rn->e[v[i] & 3];
                hcn = rn->hc;
            }
            de_edge_flux_mixed<SUBSTEP>(p, bc, gf, c, hcd, hc, i, v[i], own.e[i], er, hcn, sw, sx,
                                        sy, bflux, dtmin);
        }
        const double inv_area = 1.0 / m.area[k];
        sw *= inv_area;
        sx *= inv_area;
        sy *= inv_area;
        if (SUBSTEP == 0) {
            exw[k] = sw;
            exuh[k] = sx;
            exvh[k] = sy;
        } else {
            const DeQ r = de_advance(c, eta[k], sw, sx, sy, dt, p, err, nneg);
            out[k] = de_rk_combine<SUBSTEP>(r, bk[k]);
        }
    }

    dtmin = de_wave_min(dtmin);
    bflux = de_wave_sum(bflux);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        red_dt[wv] = dtmin;
        red_bf[wv] = bflux;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a2 = red_dt[0], b2 = red_bf[0];
        for (int i = 1; i < DE_BLOCK / 64; i++) {
            a2 = fmin(a2, red_dt[i]);
            b2 += red_bf[i];
        }
        part_dt[lb] = a2;
        part_bf[lb] = b2;
    }
}

// Tiled fused extrapolation + flux on the mixed path (any SUBSTEP): as
// de_xflux, with the FP32 edge records of de_extrapolate<float> and the FP64
// centroid depths in LDS, and the edge terms of de_flux_mixed, so the
// results are the bits of the separate mixed kernels. DE_MIXED_WAVES: the
// waves per SIMD the register allocation aims at (the autotuner's choice;
// 5, with 96 VGPRs and a 12 B spill, ran 4% faster than 4 with 100 VGPRs).
#ifndef DE_MIXED_WAVES
#define DE_MIXED_WAVES 5
#endif
template <int SUBSTEP>
__global__ void __launch_bounds__(DE_BLOCK, DE_MIXED_WAVES)
de_xflux_mixed(DeMesh m, DeParams p, const DeQ *__restrict__ q, const double *__restrict__ eta,
               DeTiles tl, DeBoundary bc, unsigned char *__restrict__ flags,
               double *__restrict__ exw, double *__restrict__ exuh, double *__restrict__ exvh,
               const DeQ *bk, DeQ *out, DeControl *ctl, double *__restrict__ part_dt,
               double *__restrict__ part_bf, double *__restrict__ part_mass,
               unsigned long long *__restrict__ nneg)
{
    // centroid depths [DE_BLOCK] tile then [hcap] halo (double); then the
    // records' components [4][3][DE_BLOCK] tile and [4][3][hcap] halo (float)
    extern __shared__ double lds[];
    __shared__ double red_dt[DE_BLOCK / 64], red_bf[DE_BLOCK / 64], red_m[DE_BLOCK / 64];
    const int lb = xcd_block();
    const int t = threadIdx.x;
    const int k0 = lb * DE_BLOCK, k = k0 + t;
    // the extrapolation's own records, requested before the control block is
    // read (bounded by the arrays' length, not yet by the scenario's end)
    int4 vk1 = make_int4(-1, -1, -1, 3);
    DeQ qk1{0.0, 0.0, 0.0, 0.0};
    if (k < m.N) {
        vk1 = m.nb[k];
        qk1 = q[k];
    }
    DeControl *cb = de_ctl(ctl, m.scen, lb);
    if (cb->done) return;
    int *err = &cb->status;
    const double dt = SUBSTEP ? cb->dt : 0.0;
    const int kend = cb->kend;
    const int hcap = tl.hcap;
    double *const hct = lds, *const hch = lds + DE_BLOCK;
    float *const ft = reinterpret_cast<float *>(lds + DE_BLOCK + hcap);
    float *const fh = ft + 12 * DE_BLOCK;
    double mass = 0.0, dtmin = 1.0e100, bflux = 0.0;
    unsigned fk = 0;
    // a halo thread's first halo triangle, loaded now so that its latency
    // hides behind the tile's extrapolation
    const int h0 = tl.start[lb], nh = tl.start[lb + 1] - h0;
    const int kh0 = t < nh ? tl.halo[h0 + t] : 0;

    if (k < kend) {
        DeExtrap<float> x;
        de_extrap_cell<float>(m, p, q, k, vk1, qk1, DeVerts{}, err, x);
        mass = x.mass;
        fk = x.flag;
        flags[k] = (unsigned char)fk;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const float4 r = de_edge32(x, i);
            ft[(0 + i) * DE_BLOCK + t] = r.x;
            ft[(3 + i) * DE_BLOCK + t] = r.y;
            ft[(6 + i) * DE_BLOCK + t] = r.z;
            ft[(9 + i) * DE_BLOCK + t] = r.w;
        }
        hct[t] = x.c.h;
    }
    for (int j = t; j < nh; j += DE_BLOCK) {
        DeExtrap<float> x;
        de_extrap_cell<float>(m, p, q, j == t ? kh0 : tl.halo[h0 + j], err, x);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const float4 r = de_edge32(x, i);
            fh[(0 + i) * hcap + j] = r.x;
            fh[(3 + i) * hcap + j] = r.y;
            fh[(6 + i) * hcap + j] = r.z;
            fh[(9 + i) * hcap + j] = r.w;
        }
        hch[j] = x.c.h;
    }
    // the flux's own records, requested before the barrier; the whole
    // geometry record is held in registers (the edge loop runs faster with
    // it there than through a pointer)
    DeQ qk2{0.0, 0.0, 0.0, 0.0};
    int4 vk2 = make_int4(-1, -1, -1, 3);
    unsigned sl2 = 0;
    DeGeoF32 gfr{};
    if (k < kend) {
        qk2 = q[k];
        vk2 = m.nb[k];
        sl2 = tl.slot[k];
        gfr = static_cast<const DeGeoF32 *>(m.gf)[k];
    }
    __syncthreads();

    if (k < kend) {
        const DeCentroid c = de_centroid(qk2, fk);
        const int4 vk = vk2;
        const int v[3] = {vk.x, vk.y, vk.z};
        const unsigned sl = sl2;
        const DeGeoF32 *gf = &gfr;
        const double hcd = hct[t];
        const float hc = float(hcd);
        double sw = 0.0, sx = 0.0, sy = 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const float4 el = make_float4(ft[(0 + i) * DE_BLOCK + t], ft[(3 + i) * DE_BLOCK + t],
                                          ft[(6 + i) * DE_BLOCK + t], ft[(9 + i) * DE_BLOCK + t]);
            float4 er = make_float4(0.f, 0.f, 0.f, 0.f);
            double hcn = 0.0;
            if (v[i] >= 0) {
                const unsigned s = (sl >> (10 * i)) & 1023u;
                const float *src;
                int stride;
                if (s == 1023u) {
                    const int tn = (v[i] >> 2) - k0;
                    src = ft + (v[i] & 3) * DE_BLOCK + tn;
                    stride = 3 * DE_BLOCK;
                    hcn = hct[tn];
                } else {
                    src = fh + (v[i] & 3) * hcap + s;
                    stride = 3 * hcap;
                    hcn = hch[s];
                }
                er = make_float4(src[0], src[stride], src[2 * stride], src[3 * stride]);
            }
            de_edge_flux_mixed<SUBSTEP>(p, bc, gf, c, hcd, hc, i, v[i], el, er, hcn, sw, sx, sy,
                                        bflux, dtmin);
        }
        const double inv_area = 1.0 / m.area[k];
        sw *= inv_area;
        sx *= inv_area;
        sy *= inv_area;
        if (SUBSTEP == 0) {
            exw[k] = sw;
            exuh[k] = sx;
            exvh[k] = sy;
        } else {
            const DeQ r = de_advance(c, eta[k], sw, sx, sy, dt, p, err, nneg);
            out[k] = de_rk_combine<SUBSTEP>(r, bk[k]);
        }
    }

    de_block_partials<SUBSTEP>(dtmin, bflux, mass, red_dt, red_bf, red_m, part_dt, part_bf,
                               part_mass, lb);
}

// friction + update + fix_negative_cells from the explicit updates of substep 0;
// `out` may alias `q` (Euler, in place: each thread reads then writes its own record)
template <bool SLOPED = false>
__global__ void __launch_bounds__(DE_BLOCK)
de_update(const int *__restrict__ scen, DeParams p, const DeQ *q, const double *__restrict__ eta,
          const unsigned char *__restrict__ flags, const double *__restrict__ exw,
          const double *__restrict__ exuh, const double *__restrict__ exvh, DeControl *ctl,
          DeQ *out, unsigned long long *__restrict__ nneg, const double *__restrict__ E = nullptr,
          int N = 0, const double *__restrict__ vx = nullptr)
{
    const int lb = xcd_block();
    DeControl *cb = de_ctl(ctl, scen, lb);
    if (cb->done) return;
    const int k = lb * DE_BLOCK + threadIdx.x;
    if (k >= cb->kend) return;
    const double dt = cb->dt;
    int *err = &cb->status;
    const DeCentroid c = de_centroid(q[k], flags[k]);
    DeSlope sl{false, 0.0, 0.0};
    if constexpr (SLOPED) sl = de_sloped_bed(E, N, k, vx);
    out[k] = de_advance(c, eta[k], exw[k], exuh[k], exvh[k], dt, p, err, nneg, sl);
}

// ---------------------------------------------------------------------------
// per-operation kernels: one kernel per call the reference's Domain makes in
// evolve_one_{euler,rk2,rk3}_step (generic_domain.py:1914-2179), over the
// centroid state q as distribute_to_vertices_and_edges left it. One
// scenario, triangles [0, kend).
// ---------------------------------------------------------------------------
// backup_centroid_values of stage, xmomentum, ymomentum (quantity.c:735-749)
__global__ void __launch_bounds__(DE_BLOCK)
de_backup(int kend, const DeQ *__restrict__ q, DeQ *__restrict__ bk)
{
    const int k = blockIdx.x * DE_BLOCK + threadIdx.x;
    if (k < kend) bk[k] = q[k];
}

// saxpy_centroid_values (quantity.c:752-769): q = a q + b backup
__global__ void __launch_bounds__(DE_BLOCK)
de_saxpy(int kend, double a, double b, DeQ *__restrict__ q, const DeQ *__restrict__ bk)
{
    const int k = blockIdx.x * DE_BLOCK + threadIdx.x;
    if (k >= kend) return;
    DeQ r = q[k];
    const DeQ s = bk[k];
    r.w = a * r.w + b * s.w;
    r.uh = a * r.uh + b * s.uh;
    r.vh = a * r.vh + b * s.vh;
    q[k] = r;
}

// rk3's closing division of the conserved centroids (generic_domain.py:2167-2170)
__global__ void __launch_bounds__(DE_BLOCK)
de_divide(int kend, double divisor, DeQ *__restrict__ q)
{
    const int k = blockIdx.x * DE_BLOCK + threadIdx.x;
    if (k >= kend) return;
    DeQ r = q[k];
    r.w = r.w / divisor;
    r.uh = r.uh / divisor;
    r.vh = r.vh / divisor;
    q[k] = r;
}

// compute_forcing_terms: Manning friction added to the momenta's
// semi-implicit updates (manning_friction_implicit, friction.py:22-76). The
// sloped form reads the vertex beds of the last extrapolation (E).
template <bool SLOPED>
__global__ void __launch_bounds__(DE_BLOCK)
de_friction(int kend, DeParams p, const DeQ *__restrict__ q, const double *__restrict__ eta,
            double *__restrict__ si_uh, double *__restrict__ si_vh,
            const double *__restrict__ E = nullptr, int N = 0,
            const double *__restrict__ vx = nullptr)
{
    const int k = blockIdx.x * DE_BLOCK + threadIdx.x;
    if (k >= kend) return;
    const DeQ r = q[k];
    const DeCentroid c{r.w, r.uh, r.vh, r.z};
    DeSlope sl{false, 0.0, 0.0};
    if constexpr (SLOPED) sl = de_sloped_bed(E, N, k, vx);
    double S;
    if (de_manning_S(c, eta[k], p, sl, S)) {
        si_uh[k] += S * c.uh;
        si_vh[k] += S * c.vh;
    }
}

// Quantity.update of stage, xmomentum and ymomentum (quantity.c:772-822): the
// explicit and semi-implicit updates with timestep dt; the semi-implicit
// updates are reset. Stage has no semi-implicit term.
__global__ void __launch_bounds__(DE_BLOCK)
de_update_centroids(int kend, DeControl *ctl, DeQ *__restrict__ q,
                    const double *__restrict__ exw, const double *__restrict__ exuh,
                    const double *__restrict__ exvh, double *__restrict__ si_uh,
                    double *__restrict__ si_vh)
{
    const int k = blockIdx.x * DE_BLOCK + threadIdx.x;
    if (k >= kend) return;
    const double dt = ctl->dt;
    int *err = &ctl->status;
    DeQ r = q[k];
    r.w = de_update_one(r.w, exw[k], 0.0, dt, err);
    r.uh = de_update_one(r.uh, exuh[k], si_uh[k], dt, err);
    r.vh = de_update_one(r.vh, exvh[k], si_vh[k], dt, err);
    q[k] = r;
    si_uh[k] = 0.0;
    si_vh[k] = 0.0;
}

// _orig_fix_negative_cells (sw_domain_orig.c:2890-2910); every triangle is full
__global__ void __launch_bounds__(DE_BLOCK)
de_fix_negative(int kend, DeQ *__restrict__ q, unsigned long long *__restrict__ nneg)
{
    const int k = blockIdx.x * DE_BLOCK + threadIdx.x;
    if (k >= kend) return;
    DeQ r = q[k];
    if (r.w - r.z < 0.0) {
        atomicAdd(nneg, 1ull);
        r.w = r.z;
        r.uh = 0.0;
        r.vh = 0.0;
        q[k] = r;
    }
}

// Fixed-order reduction of the per-block partials into the control blocks,
// in two stages so that many CUs read the partials: de_reduce_chunks splits
// each scenario's blocks into DE_RCHUNKS contiguous chunks, one 256-thread
// workgroup each; de_reduce_final adds a scenario's chunk results. MODE 0:
// flux_dt, bflux0, mass (first substep); 1: bflux1 (second substep);
// 2: mass (distribute); 3: volume; 4: bflux2 (third substep). The order of
// the additions depends only on the scenario's block count.
#define DE_RCHUNKS 32
template <int MODE>
__global__ void __launch_bounds__(256)
de_reduce_chunks(const double *__restrict__ part_dt, const double *__restrict__ part_bf,
                 const double *__restrict__ part_mass, const DeControl *__restrict__ ctl,
                 double *__restrict__ chunks)
{
    constexpr bool DT = MODE == 0, BF = MODE <= 1 || MODE == 4;
    constexpr bool MS = MODE == 0 || MODE == 2 || MODE == 3;
    __shared__ double s0[4], s1[4], s2[4];
    const int sc = blockIdx.y, c = blockIdx.x, S = gridDim.y;
    const DeControl *cb = ctl + sc;
    double a = 1.0e100, b = 0.0, m = 0.0;
    if (!((MODE < 2 || MODE == 4) && cb->done)) {
        const int b0 = cb->bbeg, b1 = cb->bend;
        const int len = (b1 - b0 + DE_RCHUNKS - 1) / DE_RCHUNKS;
        const int c0 = b0 + c * len, c1 = min(b1, c0 + len);
        for (int i = c0 + (int)threadIdx.x; i < c1; i += 256) {
            if (DT) a = fmin(a, part_dt[i]);
            if (BF) b += part_bf[i];
            if (MS) m += part_mass[i];
        }
    }
    a = de_wave_min(a);
    b = de_wave_sum(b);
    m = de_wave_sum(m);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        s0[wv] = a;
        s1[wv] = b;
        s2[wv] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; k++) {
            s0[0] = fmin(s0[0], s0[k]);
            s1[0] += s1[k];
            s2[0] += s2[k];
        }
        const size_t o = (size_t)sc * DE_RCHUNKS + c, n = (size_t)S * DE_RCHUNKS;
        chunks[o] = s0[0];
        chunks[n + o] = s1[0];
        chunks[2 * n + o] = s2[0];
    }
}

template <int MODE>
__global__ void __launch_bounds__(64)
de_reduce_final(const double *__restrict__ chunks, DeControl *ctl)
{
    const int sc = blockIdx.x, S = gridDim.x, t = threadIdx.x;
    DeControl *cb = ctl + sc;
    if ((MODE < 2 || MODE == 4) && cb->done) return;
    const size_t o = (size_t)sc * DE_RCHUNKS + t, n = (size_t)S * DE_RCHUNKS;
    double a = 1.0e100, b = 0.0, m = 0.0;
    if (t < DE_RCHUNKS) {
        a = chunks[o];
        b = chunks[n + o];
        m = chunks[2 * n + o];
    }
    a = de_wave_min(a);
    b = de_wave_sum(b);
    m = de_wave_sum(m);
    if (t == 0) {
        if (MODE == 0) {
            cb->flux_dt = a;
            cb->bflux0 = b;
            cb->mass = m;
        } else if (MODE == 1) {
            cb->bflux1 = b;
        } else if (MODE == 2) {
            cb->mass = m;
        } else if (MODE == 3) {
            cb->volume = m;
        } else {
            cb->bflux2 = b;
        }
    }
}

// one thread per scenario for the control kernels below
__device__ __forceinline__ DeControl *de_scenario(DeControl *ctl, int S)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    return s < S ? ctl + s : nullptr;
}

// boundary_flux_integral_operator.__call__ (boundary_flux_integral_operator.py:46-62)
__device__ __forceinline__ double de_flux_integral(double integral, double dt, int method,
                                                   double s0, double s1, double s2)
{
    if (method == 1) return integral + dt * s0;
    if (method == 2) return integral + 0.5 * dt * (s0 + s1);
    return integral + 1.0 / 6.0 * dt * (s0 + s1 + 4.0 * s2);
}

// Generic_Domain.update_timestep (generic_domain.py:2349-2415) on the device:
// CFL and max-timestep limit, recorded extremes, the small-step counter with
// its first-order switch and its exception (reported as DE_ERR_SMALL_TIMESTEP),
// and the clamps to the final time and the next yield time.
__global__ void de_timestep(DeControl *ctl, int S)
{
    DeControl *c = de_scenario(ctl, S);
    if (!c || c->done) return;
    const double flux_dt = c->fixed_flux_dt > 0.0 ? c->fixed_flux_dt : c->flux_dt;
    const double a = c->cfl * flux_dt;
    double ts = (c->max_dt < a) ? c->max_dt : a;           // min(CFL * flux_dt, max_dt)
    c->rec_max = (c->rec_max > ts) ? c->rec_max : ts;      // max(timestep, recorded_max)
    c->rec_min = (c->rec_min < ts) ? c->rec_min : ts;      // min(timestep, recorded_min)
    if (ts < c->min_dt) {
        c->smallsteps += 1;
        if (c->smallsteps > c->max_smallsteps) {
            c->smallsteps = 0;
            if (c->order == 1) c->status |= DE_ERR_SMALL_TIMESTEP;
            else c->order = 1;
        }
    } else {
        c->smallsteps = 0;
        if (c->order == 1 && c->default_order == 2) c->order = 2;
    }
    if (c->t + ts > c->finaltime) ts = c->finaltime - c->t;
    if (c->t + ts > c->yieldtime) ts = c->yieldtime - c->t;
    c->dt = ts;
}

// End of a step: boundary flux integral (boundary_flux_integral_operator.py:39-49),
// relative_time += dt, timestep history (one scenario only), and the yield /
// final-time test of _evolve_base (generic_domain.py:1855-1895) that turns the
// scenario's next steps into no-ops.
__global__ void de_finish_step(DeControl *ctl, int S, double *hist)
{
    DeControl *c = de_scenario(ctl, S);
    if (!c || c->done) return;
    const double dt = c->dt;
    c->integral = de_flux_integral(c->integral, dt, c->method, c->bflux0, c->bflux1, c->bflux2);
    c->t = c->t + dt;
    if (hist && c->nsteps < c->hist_cap) hist[c->nsteps] = dt;
    c->nsteps += 1;
    if (c->t >= c->finaltime - c->eps) {
        if (c->t > c->finaltime) c->status |= DE_ERR_OVERSHOOT;
        c->done = 1;
    } else if (c->t >= c->yieldtime) {
        c->done = 2;
    }
}

// host-controlled steps (one scenario): the host computed dt
__global__ void de_set_dt(DeControl *c, double dt)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) c->dt = dt;
}

// host-controlled steps: boundary flux integral after the update
__global__ void de_integrate(DeControl *c, int method)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    c->integral = de_flux_integral(c->integral, c->dt, method, c->bflux0, c->bflux1, c->bflux2);
}

// time-control parameters and state for a run of device-controlled steps
struct DeBatchSetup {
    double t, rec_min, rec_max, cfl, max_dt, min_dt, finaltime, yieldtime, fixed_flux_dt, eps;
    int smallsteps, order, default_order, max_smallsteps, hist_cap, method;
    int stop_reached;  // a scenario already at its final or yield time takes no step
};

__global__ void de_setup_batch(DeControl *ctl, int S, const DeBatchSetup *__restrict__ setup)
{
    DeControl *c = de_scenario(ctl, S);
    if (!c) return;
    const DeBatchSetup b = setup[c - ctl];
    c->t = b.t;
    c->rec_min = b.rec_min;
    c->rec_max = b.rec_max;
    c->cfl = b.cfl;
    c->max_dt = b.max_dt;
    c->min_dt = b.min_dt;
    c->finaltime = b.finaltime;
    c->yieldtime = b.yieldtime;
    c->fixed_flux_dt = b.fixed_flux_dt;
    c->eps = b.eps;
    c->smallsteps = b.smallsteps;
    c->order = b.order;
    c->default_order = b.default_order;
    c->max_smallsteps = b.max_smallsteps;
    c->hist_cap = b.hist_cap;
    c->method = b.method;
    c->nsteps = 0;
    c->gstep = 0;
    c->done = 0;
    if (b.stop_reached) {
        if (c->t >= c->finaltime - c->eps) c->done = 1;
        else if (c->t >= c->yieldtime) c->done = 2;
    }
}

// leave device-controlled mode: steps are live again
__global__ void de_clear_done(DeControl *ctl, int S)
{
    DeControl *c = de_scenario(ctl, S);
    if (c) c->done = 0;
}

// clear the error bits of every scenario (a kernel, so that no memset node
// is needed in code a caller may capture)
__global__ void de_clear_status(DeControl *ctl, int S)
{
    DeControl *c = de_scenario(ctl, S);
    if (c) c->status = 0;
}

// the per-block partial of a per-thread value: wave sums, then the block's
// waves in order (the fixed order every volume reduction uses)
__device__ __forceinline__ void de_block_sum(double v, double *__restrict__ part, int lb)
{
    __shared__ double red[DE_BLOCK / 64];
    v = de_wave_sum(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = red[0];
        for (int i = 1; i < DE_BLOCK / 64; i++) t += red[i];
        part[lb] = t;
    }
}

// per-block partial of sum(area * max(w - z, 0)) over the current state
__global__ void __launch_bounds__(DE_BLOCK)
de_volume(const int *__restrict__ scen, const DeControl *ctl, const DeQ *__restrict__ q,
          const double *__restrict__ area, double *__restrict__ part)
{
    const int lb = xcd_block();
    const DeControl *cb = scen ? ctl + scen[lb] : ctl;
    const int k = lb * DE_BLOCK + threadIdx.x;
    double v = 0.0;
    if (k < cb->kend) {
        const DeQ r = q[k];
        v = area[k] * de_height(r.w, r.z);
    }
    de_block_sum(v, part, lb);
}

// the same partials over separate stage, bed and area arrays of N triangles
__global__ void __launch_bounds__(DE_BLOCK)
de_volume_arrays(int N, const double *__restrict__ w, const double *__restrict__ z,
                 const double *__restrict__ area, double *__restrict__ part)
{
    const int lb = xcd_block();
    const int k = lb * DE_BLOCK + threadIdx.x;
    double v = 0.0;
    if (k < N) v = area[k] * de_height(w[k], z[k]);
    de_block_sum(v, part, lb);
}

// a one-scenario control block over triangles [0, N) in nblocks blocks
__global__ void de_control_init(DeControl *c, int N, int nblocks)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    c->kbeg = 0;
    c->kend = N;
    c->bbeg = 0;
    c->bend = nblocks;
    c->done = 0;
    c->status = 0;
    c->volume = 0.0;
}

__global__ void de_copy_volume(const DeControl *c, double *out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = c->volume;
}

// mixed path: FP32 bed offsets of the neighbours' centroids from the own one
// (the local datum of the mixed flux), refreshed whenever the bed changes
__global__ void __launch_bounds__(DE_BLOCK)
de_bed_offsets(int kbeg, int kend, const int4 *__restrict__ nb, const DeQ *__restrict__ q,
               DeGeoF32 *gf)
{
    const int k = kbeg + blockIdx.x * DE_BLOCK + threadIdx.x;
    if (k >= kend) return;
    const int4 v = nb[k];
    const int vv[3] = {v.x, v.y, v.z};
    const double z = q[k].z;
#pragma unroll
    for (int i = 0; i < 3; i++)
        gf[k].dzc[i] = vv[i] >= 0 ? float(q[vv[i] >> 2].z - z) : 0.f;
}

// ---------------------------------------------------------------------------
// output: gauge rows and SWW frames
// ---------------------------------------------------------------------------
// One row of gauge samples: the centroid state {w, uh, vh, z} of each gauge
// triangle and the relative time. Device-controlled steps record once per
// completed step (nsteps moved since the last row); FORCE records now with the
// host's time t (host-sequenced steps). One workgroup: the row counter moves
// after every lane has written its gauges.
template <bool FORCE>
__global__ void __launch_bounds__(256)
de_sample_gauges(DeControl *c, int G, const int *__restrict__ tri, const DeQ *__restrict__ S,
                 double t, double *__restrict__ times, double *__restrict__ vals)
{
    const long long ns = c->nsteps;
    const int row = c->grows;
    const bool take = (FORCE || ns != c->gstep) && row < c->gcap;
    if (take)
        for (int g = threadIdx.x; g < G; g += blockDim.x) {
            const DeQ q = S[tri[g]];
            double *o = vals + ((size_t)row * G + g) * 4;
            o[0] = q.w;
            o[1] = q.uh;
            o[2] = q.vh;
            o[3] = q.z;
        }
    __syncthreads();
    if (take && threadIdx.x == 0) {
        times[row] = FORCE ? t : c->t;
        c->grows = row + 1;
        c->gstep = ns;
    }
}

// gauge rows: mode 0 starts recording (cap rows, the current step is not a
// new one); mode 1 marks the rows downloaded
__global__ void de_gauge_control(DeControl *c, int mode, int cap)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (mode == 0) {
        c->gstep = c->nsteps;
        c->gcap = cap;
    }
    c->grows = 0;
}

// Node values by centroid averaging, _average_centroid_values
// (abstract_2d_finite_volumes/quantity.c:872-914): the centroid values of
// the triangles around each node summed in the node's entry order
// (vertex_value_indices), divided by their count; orphan nodes 0. Stage,
// momenta and bed together, then the storable-depth selection of
// SWW_file.store_timestep (anuga/file/sww.py:318-356): where the node depth
// w - z is below minimum_storable_height the stage is the bed and the
// momenta 0. Written as float32 [3][nodes] (the file's precision).
__global__ void __launch_bounds__(DE_BLOCK)
de_sww_nodes(int nodes, const int *__restrict__ start, const int *__restrict__ tri,
             const DeQ *__restrict__ S, double min_storable, float *__restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nodes) return;
    const int b = start[j], e = start[j + 1];
    double w = 0.0, uh = 0.0, vh = 0.0, z = 0.0;
    if (e > b) {
        for (int i = b; i < e; i++) {
            const DeQ q = S[tri[i]];
            w += q.w;
            uh += q.uh;
            vh += q.vh;
            z += q.z;
        }
        const double k = (double)(e - b);
        w = w / k;
        uh = uh / k;
        vh = vh / k;
        z = z / k;
    }
    const bool storable = w - z >= min_storable;
    out[j] = (float)(storable ? w : z);
    out[nodes + j] = storable ? (float)uh : 0.0f;
    out[2 * (size_t)nodes + j] = storable ? (float)vh : 0.0f;
}

// node values of one per-triangle quantity v[k * stride] by centroid averaging,
// as float32
__global__ void __launch_bounds__(DE_BLOCK)
de_node_values(int nodes, const int *__restrict__ start, const int *__restrict__ tri,
               const double *__restrict__ v, int stride, float *__restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nodes) return;
    const int b = start[j], e = start[j + 1];
    double a = 0.0;
    if (e > b) {
        for (int i = b; i < e; i++) a += v[(size_t)tri[i] * stride];
        a = a / (double)(e - b);
    }
    out[j] = (float)a;
}

// centroid values in the caller's order as float32: out[q][k] for the
// quantities at offsets off[q] of a record of `stride` doubles
__global__ void __launch_bounds__(DE_BLOCK)
de_gather_f32(int N, const int *__restrict__ iperm, const double *__restrict__ v, int stride,
              int nq, int4 off, float *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const double *r = v + (size_t)iperm[k] * stride;
    const int o[4] = {off.x, off.y, off.z, off.w};
    for (int q = 0; q < nq; q++) out[(size_t)q * N + k] = (float)r[o[q]];
}

// ---------------------------------------------------------------------------
// fractional-step operators
// ---------------------------------------------------------------------------
// Rate_operator.__call__ (anuga/operators/rate_operators.py:149-260) on its
// triangles after a step: local rate = factor * dt * rate; with all rates
// >= 0 the stage rises by it; otherwise (`careful`) the loss is capped at the
// depth and the momenta scale with the depth that remains. The inflow
// (local rate * area) is accumulated for the volume balance (FP64 atomics:
// a diagnostic, summed in no fixed order).
// Set_stage_operator after a step (set_quantity.py:95-131, quantity
// 'stage'): the stage centroid values of the region become the given values
// (value[i], or value0 for all); momenta and bed unchanged.
__global__ void __launch_bounds__(DE_BLOCK)
de_set_stage(const DeControl *ctl, int n, const int *__restrict__ idx,
             const double *__restrict__ value, double value0, DeQ *S)
{
    if (ctl->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    S[idx[i]].w = value ? value[i] : value0;
}

__global__ void __launch_bounds__(DE_BLOCK)
de_rate(const DeControl *ctl, int n, const int *__restrict__ idx, const double *__restrict__ rate,
        double rate0, double factor, int careful, DeQ *S, const double *__restrict__ area,
        double *influx)
{
    if (ctl->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = idx[i];
    const double r = rate ? rate[i] : rate0;
    const double dt = ctl->dt;
    DeQ q = S[k];
    double lr = factor * dt * r;
    if (!careful) {
        q.w = q.w + lr;
    } else {
        const double h = q.w - q.z;
        lr = fmax(lr, -h);
        const double lf = lr < 0.0 ? (lr + h) / (h + 1.0e-10) : 1.0;
        q.w = q.w + lr;
        q.uh = q.uh * lf;
        q.vh = q.vh * lf;
    }
    S[k] = q;
    atomicAdd(influx, lr * area[k]);
}

// ---------------------------------------------------------------------------
// Inlet_operator (anuga/structures/inlet_operator.py:75-183) on the device:
// one workgroup levels the region with the reference's numpy arithmetic
// ---------------------------------------------------------------------------
#define DE_INLET_MAX 2048
#define DE_INLET_THREADS 1024

// numpy's sum of a contiguous float64 array (pairwise_sum of
// numpy/_core/src/umath/loops_utils.h.src): fewer than 8 in order from 0.0;
// up to 128 in eight interleaved partial sums, combined as a tree, then the
// rest in order; above 128 the halves at a multiple of 8, added.
__device__ __noinline__ double de_np_leaf(const double *x, int n)
{
    if (n < 8) {
        double r = 0.0;
        for (int i = 0; i < n; i++) r += x[i];
        return r;
    }
    double r[8];
#pragma unroll
    for (int j = 0; j < 8; j++) r[j] = x[j];
    int i = 8;
    for (; i < n - (n % 8); i += 8) {
#pragma unroll
        for (int j = 0; j < 8; j++) r[j] += x[i + j];
    }
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; i++) res += x[i];
    return res;
}

template <int D>
__device__ double de_np_sum(const double *x, int n)
{
    if (n <= 128) return de_np_leaf(x, n);
    int n2 = n / 2;
    n2 -= n2 % 8;
    return de_np_sum<D - 1>(x, n2) + de_np_sum<D - 1>(x + n2, n - n2);
}
template <>
__device__ double de_np_sum<0>(const double *x, int n)
{
    return de_np_leaf(x, n);
}

// Inlet parameters: the discharge over the step is 0.5 (Q1 + Q2) (Q at the
// operator's time and one timestep later); vmode 0: momenta from the
// region's velocities, 1: from (u, v), 2: zero. stats: cumulative volume,
// total applied volume, total requested volume, applied discharge.
struct DeInlet {
    int n;
    const int *idx;       // [n] internal triangles, in the operator's order
    const double *area;   // [n]
    double total_area;    // np.sum(area)
    double Q1, Q2;
    int vmode;
    double u, v;
    double *stats;        // [4]
};

__global__ void __launch_bounds__(DE_INLET_THREADS)
de_inlet(const DeControl *ctl, DeInlet in, DeQ *S)
{
    __shared__ double st[DE_INLET_MAX], el[DE_INLET_MAX], buf[DE_INLET_MAX];
    __shared__ int order[DE_INLET_MAX];
    __shared__ double level;
    __shared__ int how, last;
    if (ctl->done) return;
    const int n = in.n, t = threadIdx.x;
    const double vp = 1.0e-6;   // anuga/config.py velocity_protection
    for (int i = t; i < n; i += DE_INLET_THREADS) {
        const DeQ q = S[in.idx[i]];
        st[i] = q.w;
        el[i] = q.z;
        buf[i] = (q.w - q.z) * in.area[i];
    }
    __syncthreads();
    const double dt = ctl->dt;
    // the step's volume (inlet_operator.py:86-96); the case is uniform
    const double Q = 0.5 * (in.Q1 + in.Q2);
    const double volume = Q * dt;
    if (t == 0) {
        const double cv = de_np_sum<6>(buf, n);   // np.sum(depths * areas)
        how = volume >= 0.0 ? 0 : (cv + volume >= 0.0 ? 1 : 2);
        level = cv;
        if (how == 0) {
            in.stats[0] += volume;     // fractional_step_volume_integral
            in.stats[2] += volume;     // total_requested_volume
            in.stats[1] += volume;     // total_applied_volume
            in.stats[3] = Q;
        } else if (how == 1) {
            in.stats[2] += volume;
            in.stats[0] += volume;
            in.stats[1] += volume;
            in.stats[3] = Q;
        } else {
            in.stats[2] += volume;
            in.stats[0] -= cv;
            in.stats[1] += -cv;
            in.stats[3] = -cv / dt;
        }
    }
    __syncthreads();
    if (how == 0) {
        // set_stages_evenly (inlet.py:192-225): stages.argsort(), ties by
        // position; then the cumulative areas and volumes in that order
        int np2 = 1;
        while (np2 < n) np2 <<= 1;
        for (int i = t; i < np2; i += DE_INLET_THREADS) {
            order[i] = i < n ? i : -1;
            if (i < n) buf[i] = in.area[i];
        }
        __syncthreads();
        for (int size = 2; size <= np2; size <<= 1) {
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int i = t; i < np2; i += DE_INLET_THREADS) {
                    const int j = i ^ stride;
                    if (j > i) {
                        const int a = order[i], b = order[j];
                        // -1 (padding) sorts last
                        const bool gt = a < 0 ? b >= 0
                                      : (b >= 0 && (st[a] > st[b] || (st[a] == st[b] && a > b)));
                        const bool up = (i & size) == 0;
                        if (gt == up) {
                            order[i] = b;
                            order[j] = a;
                        }
                    }
                }
                __syncthreads();
            }
        }
        if (t == 0) {
            double sa = buf[order[0]], sv = 0.0, sa_at = sa, sv_at = 0.0;
            int index = 0;
            for (int i = 1; i < n; i++) {
                const double nsv = sv + sa * (st[order[i]] - st[order[i - 1]]);
                sa = sa + buf[order[i]];
                sv = nsv;
                if (!(sv <= volume)) break;
                index = i;
                sa_at = sa;
                sv_at = sv;
            }
            const double depth = (volume - sv_at) / sa_at;
            level = st[order[index]] + depth;
            last = index;
        }
        __syncthreads();
        for (int i = t; i <= last; i += DE_INLET_THREADS) st[order[i]] = level;
    } else if (how == 1) {
        const double depth = (level + volume) / in.total_area;
        for (int i = t; i < n; i += DE_INLET_THREADS) st[i] = el[i] + depth;
    } else {
        for (int i = t; i < n; i += DE_INLET_THREADS) st[i] = el[i] + 0.0;
    }
    __syncthreads();
    // momenta from the velocities before the change (inlet.py:122-128)
    for (int i = t; i < n; i += DE_INLET_THREADS) {
        const int k = in.idx[i];
        DeQ q = S[k];
        const double d0 = q.w - q.z, d1 = st[i] - el[i];
        if (how == 2 || in.vmode == 2) {
            q.uh = 0.0;
            q.vh = 0.0;
        } else if (in.vmode == 1) {
            q.uh = d1 * in.u;
            q.vh = d1 * in.v;
        } else {
            const double u = q.uh * d0 / (d0 * d0 + vp);
            const double v = q.vh * d0 / (d0 * d0 + vp);
            q.uh = d1 * u;
            q.vh = d1 * v;
        }
        q.w = st[i];
        S[k] = q;
    }
}

// state of a subset of triangles (internal ids) to [n][4] {w, uh, vh, z}
__global__ void __launch_bounds__(DE_BLOCK)
de_gather_state(int n, const int *__restrict__ idx, const DeQ *__restrict__ S, DeQ *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = S[idx[i]];
}

// w, uh, vh of a subset of triangles from [n][4] (the bed is kept)
__global__ void __launch_bounds__(DE_BLOCK)
de_scatter_state(int n, const int *__restrict__ idx, const DeQ *__restrict__ in, DeQ *S)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    DeQ q = S[idx[i]];
    const DeQ v = in[i];
    q.w = v.w;
    q.uh = v.uh;
    q.vh = v.vh;
    S[idx[i]] = q;
}

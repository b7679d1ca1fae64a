// Cell-level math of the FP64 DE step in generic form, for the discrete
// adjoint.
//
// The production kernels are written for speed (packed loads, named copies,
// precision templates). Here the same arithmetic, in the same operation
// order, is restated once more over a scalar type T, so that it can run with
// T = double (the value; bit-identical to the production kernels, checked
// against the oracle on the CPU) and with T = Dual<K> (forward-mode
// derivatives along K tangent directions). The adjoint kernels evaluate a
// cell with dual numbers seeded on one group of inputs at a time and
// contract the output tangents with the output adjoints: a vector-Jacobian
// product per cell, exact for the piecewise-smooth function the forward
// computes (branches are taken on the value part, as the forward takes them).
//
// Functions are __host__ __device__ so that tests can run them on the CPU.
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ADJ_HD __host__ __device__ __forceinline__
#else
#define ADJ_HD inline
#endif

#include "de_pow.h"

// ---------------------------------------------------------------------------
// forward-mode dual numbers
// ---------------------------------------------------------------------------
template <int K>
struct Dual {
    double v;
    double d[K];
    ADJ_HD Dual() {}
    ADJ_HD Dual(double x) : v(x)
    {
#pragma unroll
        for (int k = 0; k < K; k++) d[k] = 0.0;
    }
};

template <int K>
ADJ_HD Dual<K> operator+(const Dual<K> &a, const Dual<K> &b)
{
    Dual<K> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < K; k++) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int K>
ADJ_HD Dual<K> operator-(const Dual<K> &a, const Dual<K> &b)
{
    Dual<K> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < K; k++) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int K>
ADJ_HD Dual<K> operator-(const Dual<K> &a)
{
    Dual<K> r;
    r.v = -a.v;
#pragma unroll
    for (int k = 0; k < K; k++) r.d[k] = -a.d[k];
    return r;
}
template <int K>
ADJ_HD Dual<K> operator*(const Dual<K> &a, const Dual<K> &b)
{
    Dual<K> r;
    r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < K; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int K>
ADJ_HD Dual<K> operator/(const Dual<K> &a, const Dual<K> &b)
{
    Dual<K> r;
    r.v = a.v / b.v;
    const double ib = 1.0 / b.v;
#pragma unroll
    for (int k = 0; k < K; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
template <int K> ADJ_HD Dual<K> operator+(const Dual<K> &a, double b) { return a + Dual<K>(b); }
template <int K> ADJ_HD Dual<K> operator+(double a, const Dual<K> &b) { return Dual<K>(a) + b; }
template <int K> ADJ_HD Dual<K> operator-(const Dual<K> &a, double b) { return a - Dual<K>(b); }
template <int K> ADJ_HD Dual<K> operator-(double a, const Dual<K> &b) { return Dual<K>(a) - b; }
template <int K> ADJ_HD Dual<K> operator*(const Dual<K> &a, double b) { return a * Dual<K>(b); }
template <int K> ADJ_HD Dual<K> operator*(double a, const Dual<K> &b) { return Dual<K>(a) * b; }
template <int K> ADJ_HD Dual<K> operator/(const Dual<K> &a, double b) { return a / Dual<K>(b); }
template <int K> ADJ_HD Dual<K> operator/(double a, const Dual<K> &b) { return Dual<K>(a) / b; }
template <int K> ADJ_HD Dual<K> &operator+=(Dual<K> &a, const Dual<K> &b) { return a = a + b; }
template <int K> ADJ_HD Dual<K> &operator-=(Dual<K> &a, const Dual<K> &b) { return a = a - b; }
template <int K> ADJ_HD Dual<K> &operator*=(Dual<K> &a, const Dual<K> &b) { return a = a * b; }
template <int K> ADJ_HD bool operator<(const Dual<K> &a, const Dual<K> &b) { return a.v < b.v; }
template <int K> ADJ_HD bool operator>(const Dual<K> &a, const Dual<K> &b) { return a.v > b.v; }
template <int K> ADJ_HD bool operator<=(const Dual<K> &a, const Dual<K> &b) { return a.v <= b.v; }
template <int K> ADJ_HD bool operator>=(const Dual<K> &a, const Dual<K> &b) { return a.v >= b.v; }
template <int K> ADJ_HD bool operator==(const Dual<K> &a, const Dual<K> &b) { return a.v == b.v; }

// the value part of a scalar
ADJ_HD double adj_val(double x) { return x; }
template <int K> ADJ_HD double adj_val(const Dual<K> &x) { return x.v; }

// max / min: the argument the forward's fmax / fmin returns (ties: the first)
ADJ_HD double adj_max(double a, double b) { return fmax(a, b); }
ADJ_HD double adj_min(double a, double b) { return fmin(a, b); }
template <int K> ADJ_HD Dual<K> adj_max(const Dual<K> &a, const Dual<K> &b)
{
    return (a.v >= b.v) ? a : b;
}
template <int K> ADJ_HD Dual<K> adj_min(const Dual<K> &a, const Dual<K> &b)
{
    return (a.v <= b.v) ? a : b;
}

// a ? x : (b ? y : w). For duals the choice is an arithmetic blend with 0/1
// weights: a select between elements of local arrays compiles to a load from a
// selected address, which keeps the arrays in scratch memory. The blend's
// value equals the chosen one for finite inputs (up to the sign of a zero);
// the double versions, which must reproduce the forward, select exactly.
ADJ_HD double adj_sel(bool a, double x, double y) { return a ? x : y; }
template <int K>
ADJ_HD Dual<K> adj_sel(bool a, const Dual<K> &x, const Dual<K> &y)
{
    const double ma = a ? 1.0 : 0.0, mb = 1.0 - ma;
    Dual<K> r;
    r.v = ma * x.v + mb * y.v;
#pragma unroll
    for (int k = 0; k < K; k++) r.d[k] = ma * x.d[k] + mb * y.d[k];
    return r;
}
ADJ_HD double adj_sel3(bool a, bool b, double x, double y, double w)
{
    return a ? x : (b ? y : w);
}
template <int K>
ADJ_HD Dual<K> adj_sel3(bool a, bool b, const Dual<K> &x, const Dual<K> &y, const Dual<K> &w)
{
    const double ma = a ? 1.0 : 0.0, mb = (!a && b) ? 1.0 : 0.0, mc = 1.0 - ma - mb;
    Dual<K> r;
    r.v = ma * x.v + mb * y.v + mc * w.v;
#pragma unroll
    for (int k = 0; k < K; k++) r.d[k] = ma * x.d[k] + mb * y.d[k] + mc * w.d[k];
    return r;
}

// the value without derivative
ADJ_HD double adj_const(double a) { return a; }
template <int K> ADJ_HD Dual<K> adj_const(const Dual<K> &a) { return Dual<K>(a.v); }

ADJ_HD double adj_sqrt(double a) { return sqrt(a); }
template <int K> ADJ_HD Dual<K> adj_sqrt(const Dual<K> &a)
{
    Dual<K> r;
    r.v = sqrt(a.v);
    // sqrt is not differentiable at 0; every use here is sqrt(g h) or a
    // momentum magnitude whose product with the tangent direction vanishes there
    const double s = r.v > 0.0 ? 0.5 / r.v : 0.0;
#pragma unroll
    for (int k = 0; k < K; k++) r.d[k] = a.d[k] * s;
    return r;
}

// h^(7/3) with the correctly rounded value of the FP64 path
ADJ_HD double adj_pow73(double h) { return de_pow_7_3(h); }
template <int K> ADJ_HD Dual<K> adj_pow73(const Dual<K> &h)
{
    Dual<K> r;
    r.v = de_pow_7_3(h.v);
    const double s = (7.0 / 3.0) * r.v / h.v;
#pragma unroll
    for (int k = 0; k < K; k++) r.d[k] = h.d[k] * s;
    return r;
}

// ---------------------------------------------------------------------------
// parameters and flags (the same values the production kernels use)
// ---------------------------------------------------------------------------
struct AdjParams {
    double g, eps, mah;
    double bw, bwd, buh, buhd, bvh, bvhd;
    // limiter jumps below this size are rounding noise: the limiter's factor
    // is held at its value there (0: exact derivative everywhere)
    double freeze;
};

#define ADJ_F_CLAMP 1u
#define ADJ_F_KEEP 2u

// ---------------------------------------------------------------------------
// limiter and gradients (sw_domain_orig.c:1306-1333, :2238-2277, :2470-2504)
// ---------------------------------------------------------------------------
template <class T>
ADJ_HD void adj_limit(T d[3], T qmin, T qmax, T beta, double freeze = 0.0)
{
    T r = T(1000.0), r0 = T(1.0);
    const T tiny = T(1.0e-100);
    const T mtiny = T(-1.0e-100);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (d[i] < mtiny) r0 = qmin / d[i];
        if (d[i] > tiny) r0 = qmax / d[i];
        r = adj_min(r0, r);
    }
    T phi = adj_min(r * beta, T(1.0));
    // At jumps of rounding size (still water reached only by rounding noise)
    // the ratio qmin / d of two noise-sized numbers has derivatives of 1/d,
    // which the adjoint of many steps amplifies without bound. There the
    // factor is held at its value; the forward value is unchanged.
    if (freeze > 0.0 && fabs(adj_val(d[0])) < freeze && fabs(adj_val(d[1])) < freeze &&
        fabs(adj_val(d[2])) < freeze)
        phi = adj_const(phi);
    d[0] = d[0] * phi;
    d[1] = d[1] * phi;
    d[2] = d[2] * phi;
}

template <class T>
ADJ_HD void adj_aux_jumps(T qk, T q0, T q1, T q2, double dx1, double dx2, double dy1, double dy2,
                          double inv_area2, const double dxv[3], const double dyv[3], T beta,
                          T d[3], double freeze = 0.0)
{
    const T dq0 = q0 - qk;
    const T dq1 = q1 - q0;
    const T dq2 = q2 - q0;
    T a = T(dy2) * dq1 - T(dy1) * dq2;
    a *= T(inv_area2);
    T b = T(dx1) * dq2 - T(dx2) * dq1;
    b *= T(inv_area2);
#pragma unroll
    for (int i = 0; i < 3; i++) d[i] = a * T(dxv[i]) + b * T(dyv[i]);
    const T zero = T(0.0);
    const T qmax = adj_max(adj_max(dq0, adj_max(dq0 + dq1, dq0 + dq2)), zero);
    const T qmin = adj_min(adj_min(dq0, adj_min(dq0 + dq1, dq0 + dq2)), zero);
    adj_limit(d, qmin, qmax, beta, freeze);
}

template <class T>
ADJ_HD void adj_line_jumps(T qk, T q1, double dx2, double dy2, const double dxv[3],
                           const double dyv[3], T beta, T d[3], double freeze = 0.0)
{
    const T dq1 = q1 - qk;
    const T a = dq1 * T(dx2);
    const T b = dq1 * T(dy2);
#pragma unroll
    for (int i = 0; i < 3; i++) d[i] = a * T(dxv[i]) + b * T(dyv[i]);
    const bool up = dq1 >= T(0.0);
    const T qmin = adj_sel(up, T(0.0), dq1);
    const T qmax = adj_sel(up, dq1, T(0.0));
    adj_limit(d, qmin, qmax, beta, freeze);
}

// ---------------------------------------------------------------------------
// extrapolation of one triangle (sw_domain_orig.c:1994-2680)
// ---------------------------------------------------------------------------
// Inputs: raw state (w, uh, vh, z) of the triangle (slot 0) and of its three
// surrogate neighbours (slots 1-3; a boundary edge's surrogate is the triangle
// itself), the flags of the four (protect clamp / momentum kept, which the
// forward derived from the same values and from the neighbours' neighbours),
// the number of boundary edges, which edges are interior, and the geometry:
// edge-midpoint offsets dxv/dyv and the surrogate-centroid differences.
// Output: stage, height, xmom, ymom on the three edges.
struct AdjExtrapGeo {
    double dxv[3], dyv[3];
    double dx1, dx2, dy1, dy2;   // nbound <= 1: centroid differences sur1-sur0, sur2-sur0
    double lx, ly;               // nbound == 2: first interior surrogate - own centroid
    int nbound;
    bool interior[3];
};

template <class T>
ADJ_HD void adj_extrap_cell(const T w[4], const T uh[4], const T vh[4], const T z[4],
                            const unsigned fl[4], const AdjExtrapGeo &gm, const AdjParams &p,
                            T ew[3], T eh[3], T euh[3], T evh[3])
{
    const T zero = T(0.0);
    // own centroid after protect and the velocity conversion
    const T cw = (fl[0] & ADJ_F_CLAMP) ? z[0] : w[0];
    const T ch = adj_max(cw - z[0], zero);
    T cu = zero, cv = zero;
    if (fl[0] & ADJ_F_KEEP) {
        const T inv = T(1.0) / ch;
        cu = uh[0] * inv;
        cv = vh[0] * inv;
    }
    // surrogates
    T nw[3], nh[3], nu[3], nv[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const T hs = adj_max(w[i + 1] - z[i + 1], zero);
        nw[i] = (fl[i + 1] & ADJ_F_CLAMP) ? z[i + 1] : w[i + 1];
        nh[i] = hs;
        nu[i] = zero;
        nv[i] = zero;
        if (fl[i + 1] & ADJ_F_KEEP) {
            const T inv = T(1.0) / hs;
            nu[i] = uh[i + 1] * inv;
            nv[i] = vh[i + 1] * inv;
        }
    }
    T jw[3] = {zero, zero, zero}, jh[3] = {zero, zero, zero};
    T ju[3] = {zero, zero, zero}, jv[3] = {zero, zero, zero};
    bool constant = (gm.nbound == 3);
    bool gw = true, gu = true, gv = true;
    if (gm.nbound <= 1) {
        const double area2 = gm.dy2 * gm.dx1 - gm.dy1 * gm.dx2;
        if (area2 <= 0.0) {
            constant = true;
        } else {
            const double a_tmp = 0.3, b_tmp = 0.1;
            const double c_tmp = 1.0 / (a_tmp - b_tmp);
            const double d_tmp = 1.0 - (c_tmp * a_tmp);
            const T hmin = adj_min(adj_min(nh[0], adj_min(nh[1], nh[2])), ch);
            const T hmax = adj_max(adj_max(nh[0], adj_max(nh[1], nh[2])), ch);
            T hfactor = adj_max(zero, adj_min(T(c_tmp) * adj_max(hmin, zero) / adj_max(ch, T(1.0e-06)) + T(d_tmp),
                                              adj_min(T(c_tmp) * adj_max(ch, zero) / adj_max(hmax, T(1.0e-06)) + T(d_tmp),
                                                      T(1.0))));
            hfactor = adj_min(T(1.2) * adj_max(hmin - T(p.mah), zero) / (adj_max(hmin, zero) + T(1.0) * T(p.mah)),
                              hfactor);
            const double inv_area2 = 1.0 / area2;
            T beta = T(p.bwd) + (T(p.bw) - T(p.bwd)) * hfactor;
            gw = beta > zero;
            if (gw) {
                adj_aux_jumps<T>(cw, nw[0], nw[1], nw[2], gm.dx1, gm.dx2, gm.dy1, gm.dy2,
                                 inv_area2, gm.dxv, gm.dyv, beta, jw, p.freeze);
                adj_aux_jumps<T>(ch, nh[0], nh[1], nh[2], gm.dx1, gm.dx2, gm.dy1, gm.dy2,
                                 inv_area2, gm.dxv, gm.dyv, beta, jh, p.freeze);
            }
            beta = T(p.buhd) + (T(p.buh) - T(p.buhd)) * hfactor;
            gu = beta > zero;
            if (gu)
                adj_aux_jumps<T>(cu, nu[0], nu[1], nu[2], gm.dx1, gm.dx2, gm.dy1, gm.dy2,
                                 inv_area2, gm.dxv, gm.dyv, beta, ju, p.freeze);
            beta = T(p.bvhd) + (T(p.bvh) - T(p.bvhd)) * hfactor;
            gv = beta > zero;
            if (gv)
                adj_aux_jumps<T>(cv, nv[0], nv[1], nv[2], gm.dx1, gm.dx2, gm.dy1, gm.dy2,
                                 inv_area2, gm.dxv, gm.dyv, beta, jv, p.freeze);
        }
    } else if (gm.nbound == 2) {
        const bool e0 = gm.interior[0], e1 = gm.interior[1];
        if (!(gm.interior[0] || gm.interior[1] || gm.interior[2])) {
            constant = true;
        } else {
            // the first interior surrogate, selected (not indexed) so that the
            // arrays stay in registers
            const T sw = adj_sel3(e0, e1, nw[0], nw[1], nw[2]);
            const T sh = adj_sel3(e0, e1, nh[0], nh[1], nh[2]);
            const T su = adj_sel3(e0, e1, nu[0], nu[1], nu[2]);
            const T sv = adj_sel3(e0, e1, nv[0], nv[1], nv[2]);
            const double area2 = gm.lx * gm.lx + gm.ly * gm.ly;
            double dx2 = 1.0 / area2;
            const double dy2 = dx2 * gm.ly;
            dx2 *= gm.lx;
            const T bw = T(p.bw);
            adj_line_jumps<T>(cw, sw, dx2, dy2, gm.dxv, gm.dyv, bw, jw, p.freeze);
            adj_line_jumps<T>(ch, sh, dx2, dy2, gm.dxv, gm.dyv, bw, jh, p.freeze);
            adj_line_jumps<T>(cu, su, dx2, dy2, gm.dxv, gm.dyv, bw, ju, p.freeze);
            adj_line_jumps<T>(cv, sv, dx2, dy2, gm.dxv, gm.dyv, bw, jv, p.freeze);
        }
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        ew[i] = (constant || !gw) ? cw : cw + jw[i];
        eh[i] = (constant || !gw) ? ch : ch + jh[i];
        const T eu = (constant || !gu) ? cu : cu + ju[i];
        const T ev = (constant || !gv) ? cv : cv + jv[i];
        euh[i] = eu * eh[i];
        evh[i] = ev * eh[i];
    }
}

// ---------------------------------------------------------------------------
// central-upwind flux across one edge, low_froude = 0 (sw_domain_orig.c:284-505)
// ---------------------------------------------------------------------------
template <class T>
ADJ_HD void adj_flux_central(T ql0, T ql1, T ql2, T qr0, T qr1, T qr2, T h_left, T h_right,
                             T hle, T hre, double n1, double n2, double eps, T ze, double g,
                             T ef[3], T &pressure_flux)
{
    const T zero = T(0.0);
    if (h_left == zero && h_right == zero) {
        ef[0] = ef[1] = ef[2] = zero;
        pressure_flux = zero;
        return;
    }
    const T l1 = T(n1) * ql1 + T(n2) * ql2;
    const T l2 = T(-n2) * ql1 + T(n1) * ql2;
    const T r1 = T(n1) * qr1 + T(n2) * qr2;
    const T r2 = T(-n2) * qr1 + T(n1) * qr2;
    T ul, uhl, vhl, ur, uhr, vhr;
    uhl = l1;
    vhl = l2;
    if (hle > zero) {
        const T t = T(1.0) / hle;
        ul = uhl * t;
        uhl = h_left * ul;
        vhl = h_left * t * vhl;
    } else {
        ul = uhl = vhl = zero;
    }
    uhr = r1;
    vhr = r2;
    if (hre > zero) {
        const T t = T(1.0) / hre;
        ur = uhr * t;
        uhr = h_right * ur;
        vhr = h_right * t * vhr;
    } else {
        ur = uhr = vhr = zero;
    }
    const T cl = adj_sqrt(T(g) * h_left);
    const T cr = adj_sqrt(T(g) * h_right);
    T s_max = adj_max(ul + cl, ur + cr);
    if (s_max < zero) s_max = zero;
    T s_min = adj_min(ul - cl, ur - cr);
    if (s_min > zero) s_min = zero;
    const T fl0 = ul * h_left, fl1 = ul * uhl, fl2 = ul * vhl;
    const T fr0 = ur * h_right, fr1 = ur * uhr, fr2 = ur * vhr;
    const T denom = s_max - s_min;
    if (denom < T(eps)) {
        ef[0] = ef[1] = ef[2] = zero;
        pressure_flux = T(0.5) * T(g) * T(0.5) * (h_left * h_left + h_right * h_right);
        return;
    }
    const T inv = T(1.0) / adj_max(denom, T(1.0e-100));
    const T sms = s_max * s_min;
    T e0 = s_max * fl0 - s_min * fr0;
    e0 += sms * (adj_max(qr0, ze) - adj_max(ql0, ze));
    e0 *= inv;
    T e1 = s_max * fl1 - s_min * fr1;
    e1 += T(1.0) * sms * (uhr - uhl);
    e1 *= inv;
    T e2 = s_max * fl2 - s_min * fr2;
    e2 += T(1.0) * sms * (vhr - vhl);
    e2 *= inv;
    pressure_flux = T(0.5) * T(g) * (s_max * h_left * h_left - s_min * h_right * h_right) * inv;
    ef[0] = e0;
    ef[1] = T(n1) * e1 + T(-n2) * e2;
    ef[2] = -T(-n2) * e1 + T(n1) * e2;
}

// One edge's contribution to the triangle's explicit update, before the
// division by the area (sw_domain_orig.c:879-1004 and :1118-1132), as the
// five terms the forward accumulates in this order: sw += out[0], sx += out[1],
// sy += out[2], sx -= out[3], sy -= out[4]. Inputs are
// the triangle's own edge values (left), the neighbour's edge values (right;
// for a boundary edge derived from the left ones by the boundary kind), and
// the triangle's centroid depth and bed. kind: -1 interior, else the
// boundary kind; dir: Dirichlet [stage, xmom, ymom]; cen: the centroid
// stage / xmom / ymom (centroid-transmissive kind).
template <class T>
ADJ_HD void adj_edge_update(T ql0, T hle, T ql1, T ql2, T qr0_in, T hre_in, T qr1_in, T qr2_in,
                            int kind, const double dir[3], const T cen[3], T hc, T zc, double n1,
                            double n2, double length, const AdjParams &p, T out[5])
{
    const T zero = T(0.0);
    const T zl = ql0 - hle;
    T qr0, qr1, qr2, zr, hre;
    if (kind < 0) {
        qr0 = qr0_in;
        qr1 = qr1_in;
        qr2 = qr2_in;
        hre = hre_in;
        zr = qr0 - hre;
    } else {
        if (kind == 0) {
            const T r1 = -ql1 * T(n1) - ql2 * T(n2);
            const T r2 = -ql1 * T(n2) + ql2 * T(n1);
            qr0 = ql0;
            qr1 = T(n1) * r1 - T(n2) * r2;
            qr2 = T(n2) * r1 + T(n1) * r2;
        } else if (kind == 1) {
            qr0 = ql0;
            qr1 = ql1;
            qr2 = ql2;
        } else if (kind == 2) {
            qr0 = T(dir[0]);
            qr1 = T(dir[1]);
            qr2 = T(dir[2]);
        } else {
            qr0 = cen[0];
            qr1 = cen[1];
            qr2 = cen[2];
        }
        zr = zl;
        hre = adj_max(qr0 - zr, zero);
    }
    const T z_half = adj_max(zl, zr);
    const T h_left = adj_max(hle + zl - z_half, zero);
    const T h_right = adj_max(hre + zr - z_half, zero);
    T ef[3], pflux;
    adj_flux_central<T>(ql0, ql1, ql2, qr0, qr1, qr2, h_left, h_right, hle, hre, n1, n2, p.eps,
                        z_half, p.g, ef, pflux);
    const T len = T(length);
    ef[0] = ef[0] * len;
    ef[1] = ef[1] * len;
    ef[2] = ef[2] * len;
    const T pg = len * (-T(p.g) * T(0.5) * (h_left * h_left - hle * hle - (hle + hc) * (zl - zc)) + pflux);
    out[0] = -ef[0];
    out[1] = -ef[1];
    out[2] = -ef[2];
    out[3] = T(n1) * pg;
    out[4] = T(n2) * pg;
}

// ---------------------------------------------------------------------------
// friction + update + fix_negative_cells of one triangle
// (sw_domain_orig.c:2805-2838, quantity.c:772-822, sw_domain_orig.c:2890-2910)
// ---------------------------------------------------------------------------
template <class T>
ADJ_HD T adj_update_one(T q, T ex, T si, double dt)
{
    if (q == T(0.0)) si = T(0.0);
    else si = si / q;
    q = q + T(dt) * ex;
    const T den = T(1.0) - T(dt) * si;
    return q / den;
}

// c: centroid w, uh, vh, z after protect; ex: explicit updates (already / area)
template <class T>
ADJ_HD void adj_advance_cell(T cw, T cuh, T cvh, T cz, T eta, T exw, T exuh, T exvh, double dt,
                             const AdjParams &p, T out[3])
{
    T si_uh = T(0.0), si_vh = T(0.0);
    if (eta > T(p.mah)) {
        const T h = cw - cz;
        if (h >= T(p.mah)) {
            T S = -T(p.g) * eta * eta * adj_sqrt((cuh * cuh + cvh * cvh));
            S = S / adj_pow73(h);
            si_uh = si_uh + S * cuh;
            si_vh = si_vh + S * cvh;
        }
    }
    T rw = adj_update_one<T>(cw, exw, T(0.0), dt);
    T ru = adj_update_one<T>(cuh, exuh, si_uh, dt);
    T rv = adj_update_one<T>(cvh, exvh, si_vh, dt);
    if (rw - cz < T(0.0)) {
        rw = cz;
        ru = T(0.0);
        rv = T(0.0);
    }
    out[0] = rw;
    out[1] = ru;
    out[2] = rv;
}

// ---------------------------------------------------------------------------
// per-triangle assembly over the library's device layouts
// ---------------------------------------------------------------------------
// Layouts (internal triangle order): state records S[4k..4k+3] = w, uh, vh, z;
// nb[4k+i] = (n << 2) | m for the neighbour n across edge i (its edge m), or
// -(j+1) for boundary edge j, nb[4k+3] = number of boundary edges; cen[2k..];
// gx[6k..] = edge-midpoint offsets x0..x2, y0..y2; per-edge arrays [3][N];
// edge values E = w, h, uh, vh blocks of [3][N]; adjoints of the state [3][N].
struct AdjView {
    int N, M;
    int gx_stride;        // doubles per triangle in gx (the first 6: x0..x2, y0..y2)
    const int *nb;
    const double *cen, *gx, *nx, *ny, *len, *area, *eta;
    const signed char *bkind;
    const double *bval;   // [3][M]
    AdjParams p;
    double dt;
};

// accumulate into an adjoint: several triangles add to the same entry
ADJ_HD void adj_add(double *a, double v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    atomicAdd(a, v);
#else
    *a += v;
#endif
}

ADJ_HD void adj_neighbours(const AdjView &m, int k, int v[3], int sur[3])
{
#pragma unroll
    for (int i = 0; i < 3; i++) {
        v[i] = m.nb[4 * k + i];
        sur[i] = v[i] >= 0 ? (v[i] >> 2) : k;
    }
}

ADJ_HD AdjExtrapGeo adj_geometry(const AdjView &m, int k, const int v[3], const int sur[3])
{
    AdjExtrapGeo g;
    for (int i = 0; i < 3; i++) {
        g.dxv[i] = m.gx[(size_t)m.gx_stride * k + i];
        g.dyv[i] = m.gx[(size_t)m.gx_stride * k + 3 + i];
        g.interior[i] = v[i] >= 0;
    }
    g.nbound = m.nb[4 * k + 3];
    g.dx1 = g.dx2 = g.dy1 = g.dy2 = g.lx = g.ly = 0.0;
    if (g.nbound <= 1) {
        const double *c0 = m.cen + 2 * sur[0], *c1 = m.cen + 2 * sur[1], *c2 = m.cen + 2 * sur[2];
        g.dx1 = c1[0] - c0[0];
        g.dx2 = c2[0] - c0[0];
        g.dy1 = c1[1] - c0[1];
        g.dy2 = c2[1] - c0[1];
    } else if (g.nbound == 2) {
        const int s1 = v[0] >= 0 ? sur[0] : (v[1] >= 0 ? sur[1] : sur[2]);
        g.lx = m.cen[2 * s1] - m.cen[2 * k];
        g.ly = m.cen[2 * s1 + 1] - m.cen[2 * k + 1];
    }
    return g;
}

// protect / momentum flags of triangle k from the state (the production
// extrapolation's flag byte)
ADJ_HD unsigned adj_flags(const AdjView &m, const double *S, int k)
{
    int v[3], sur[3];
    adj_neighbours(m, k, v, sur);
    const double w = S[4 * k], z = S[4 * k + 3], mah = m.p.mah;
    const double hc = w - z;
    const bool clamp = (hc < mah * 1.0) && (hc <= 0.0) && (w < z);
    const double cw = clamp ? z : w;
    const double ch = fmax(cw - z, 0.0);
    bool dry = true;
    for (int i = 0; i < 3; i++)
        dry = dry && (v[i] < 0 || fmax(S[4 * sur[i]] - S[4 * sur[i] + 3], 0.0) < mah);
    const bool keep = (ch > mah) && !dry;
    return (clamp ? ADJ_F_CLAMP : 0u) | (keep ? ADJ_F_KEEP : 0u);
}

// centroid state after the flags (what the flux and the update read)
template <class T>
ADJ_HD void adj_centroid(const T q[4], unsigned f, T c[4])
{
    c[0] = (f & ADJ_F_CLAMP) ? q[3] : q[0];
    c[1] = (f & ADJ_F_KEEP) ? q[1] : T(0.0);
    c[2] = (f & ADJ_F_KEEP) ? q[2] : T(0.0);
    c[3] = q[3];
}

// ---- forward (T = double), for the CPU harness -------------------------------
ADJ_HD void adj_fwd_extrap(const AdjView &m, const double *S, const unsigned char *fl, int k,
                           double *E)
{
    int v[3], sur[3];
    adj_neighbours(m, k, v, sur);
    const AdjExtrapGeo g = adj_geometry(m, k, v, sur);
    const int id[4] = {k, sur[0], sur[1], sur[2]};
    double w[4], uh[4], vh[4], z[4];
    unsigned f[4];
    for (int j = 0; j < 4; j++) {
        w[j] = S[4 * id[j]];
        uh[j] = S[4 * id[j] + 1];
        vh[j] = S[4 * id[j] + 2];
        z[j] = S[4 * id[j] + 3];
        f[j] = fl[id[j]];
    }
    double ew[3], eh[3], euh[3], evh[3];
    adj_extrap_cell<double>(w, uh, vh, z, f, g, m.p, ew, eh, euh, evh);
    const size_t N = m.N;
    for (int i = 0; i < 3; i++) {
        E[i * N + k] = ew[i];
        E[3 * N + i * N + k] = eh[i];
        E[6 * N + i * N + k] = euh[i];
        E[9 * N + i * N + k] = evh[i];
    }
}

// inputs of edge i of triangle k: own edge values, the neighbour's, boundary kind
ADJ_HD void adj_edge_inputs(const AdjView &m, const double *E, int k, int i, int vi, double L[4],
                            double R[4], int &kind, double dir[3])
{
    const size_t N = m.N;
    const size_t ki = (size_t)i * N + k;
    L[0] = E[ki];
    L[1] = E[3 * N + ki];
    L[2] = E[6 * N + ki];
    L[3] = E[9 * N + ki];
    dir[0] = dir[1] = dir[2] = 0.0;
    if (vi >= 0) {
        const size_t nm = (size_t)(vi & 3) * N + (vi >> 2);
        R[0] = E[nm];
        R[1] = E[3 * N + nm];
        R[2] = E[6 * N + nm];
        R[3] = E[9 * N + nm];
        kind = -1;
    } else {
        const int j = -vi - 1;
        kind = m.bkind[j];
        R[0] = R[1] = R[2] = R[3] = 0.0;
        if (kind == 2) {
            dir[0] = m.bval[j];
            dir[1] = m.bval[m.M + j];
            dir[2] = m.bval[2 * m.M + j];
        }
    }
}

// one edge's five update terms from (L, R, C = centroid w, uh, vh, z)
template <class T>
ADJ_HD void adj_edge_terms(const AdjView &m, int k, int i, const T L[4], const T R[4],
                           const T C[4], int kind, const double dir[3], T out[5])
{
    const T zero = T(0.0);
    const T hc = adj_max(C[0] - C[3], zero);
    const T cen[3] = {C[0], C[1], C[2]};
    const size_t N = m.N, ki = (size_t)i * N + k;
    // argument order of the edge function: ql0, hle, ql1, ql2, qr0, hre, qr1, qr2
    adj_edge_update<T>(L[0], L[1], L[2], L[3], R[0], R[1], R[2], R[3], kind, dir, cen, hc, C[3],
                       m.nx[ki], m.ny[ki], m.len[ki], m.p, out);
}

ADJ_HD void adj_fwd_flux(const AdjView &m, const double *S, const unsigned char *fl,
                         const double *E, int k, double ex[3])
{
    int v[3], sur[3];
    adj_neighbours(m, k, v, sur);
    double q[4] = {S[4 * k], S[4 * k + 1], S[4 * k + 2], S[4 * k + 3]}, c[4];
    adj_centroid<double>(q, fl[k], c);
    double sw = 0.0, sx = 0.0, sy = 0.0;
    for (int i = 0; i < 3; i++) {
        double L[4], R[4], dir[3], o[5];
        int kind;
        adj_edge_inputs(m, E, k, i, v[i], L, R, kind, dir);
        // the edge function takes (w, h, uh, vh) in the order ql0, hle, ql1, ql2
        adj_edge_terms<double>(m, k, i, L, R, c, kind, dir, o);
        sw += o[0];
        sx += o[1];
        sy += o[2];
        sx -= o[3];
        sy -= o[4];
    }
    const double inv_area = 1.0 / m.area[k];
    ex[0] = sw * inv_area;
    ex[1] = sx * inv_area;
    ex[2] = sy * inv_area;
}

ADJ_HD void adj_fwd_advance(const AdjView &m, const double *S, const unsigned char *fl,
                            const double ex[3], int k, double out[3])
{
    double q[4] = {S[4 * k], S[4 * k + 1], S[4 * k + 2], S[4 * k + 3]}, c[4];
    adj_centroid<double>(q, fl[k], c);
    adj_advance_cell<double>(c[0], c[1], c[2], c[3], m.eta[k], ex[0], ex[1], ex[2], m.dt, m.p, out);
}

// ---- adjoints ----------------------------------------------------------------------
// Map an adjoint of the centroid values (w, uh, vh, z after the flags) onto
// the raw state of triangle k and the bed.
ADJ_HD void adj_scatter_centroid(unsigned f, const double lc[4], double *lamS, double *zbar,
                                 int N, int k)
{
    const double lw = (f & ADJ_F_CLAMP) ? 0.0 : lc[0];
    const double lz = lc[3] + ((f & ADJ_F_CLAMP) ? lc[0] : 0.0);
    const double lu = (f & ADJ_F_KEEP) ? lc[1] : 0.0;
    const double lv = (f & ADJ_F_KEEP) ? lc[2] : 0.0;
    if (lw != 0.0) adj_add(lamS + k, lw);
    if (lu != 0.0) adj_add(lamS + N + k, lu);
    if (lv != 0.0) adj_add(lamS + 2 * (size_t)N + k, lv);
    if (lz != 0.0) adj_add(zbar + k, lz);
}

// Update of triangle k: lam_out (w, uh, vh of the result) -> adjoints of the
// raw state (lamS, zbar), of the friction coefficient (nbar) and of the
// explicit updates (lam_ex, stored).
ADJ_HD void adj_bwd_advance(const AdjView &m, const double *S, const unsigned char *fl,
                            const double *exw, const double *exuh, const double *exvh,
                            const double lam_out[3], int k, double *lamS, double *zbar,
                            double *nbar, double *lam_ex)
{
    typedef Dual<4> D;
    const int N = m.N;
    double q[4] = {S[4 * k], S[4 * k + 1], S[4 * k + 2], S[4 * k + 3]}, c[4];
    const unsigned f = fl[k];
    adj_centroid<double>(q, f, c);
    const double in[8] = {c[0], c[1], c[2], c[3], m.eta[k], exw[k], exuh[k], exvh[k]};
    double g[8];
    for (int pass = 0; pass < 2; pass++) {
        D x[8];
        for (int j = 0; j < 8; j++) {
            x[j] = D(in[j]);
            if (j / 4 == pass) x[j].d[j % 4] = 1.0;
        }
        D o[3];
        adj_advance_cell<D>(x[0], x[1], x[2], x[3], x[4], x[5], x[6], x[7], m.dt, m.p, o);
        for (int t = 0; t < 4; t++)
            g[4 * pass + t] = lam_out[0] * o[0].d[t] + lam_out[1] * o[1].d[t] + lam_out[2] * o[2].d[t];
    }
    const double lc[4] = {g[0], g[1], g[2], g[3]};
    adj_scatter_centroid(f, lc, lamS, zbar, N, k);
    if (g[4] != 0.0) adj_add(nbar + k, g[4]);
    lam_ex[k] = g[5];
    lam_ex[N + k] = g[6];
    lam_ex[2 * (size_t)N + k] = g[7];
}

// Flux of triangle k: lam_ex (its explicit update) -> adjoints of its own and
// its neighbours' edge values (lamE) and of its centroid state. KF tangents
// per pass (the four components of one input set in 4 / KF passes): a
// narrower dual lowers register pressure; the tangents are the same numbers.
template <int KF = 4>
ADJ_HD void adj_bwd_flux(const AdjView &m, const double *S, const unsigned char *fl,
                         const double *E, const double *lam_ex, int k, double *lamE,
                         double *lamS, double *zbar)
{
    typedef Dual<KF> D;
    const int N = m.N;
    int v[3], sur[3];
    adj_neighbours(m, k, v, sur);
    double q[4] = {S[4 * k], S[4 * k + 1], S[4 * k + 2], S[4 * k + 3]}, c[4];
    const unsigned f = fl[k];
    adj_centroid<double>(q, f, c);
    const double inv_area = 1.0 / m.area[k];
    const double lw = lam_ex[k] * inv_area, lx = lam_ex[N + k] * inv_area;
    const double ly = lam_ex[2 * (size_t)N + k] * inv_area;
    const double lo[5] = {lw, lx, ly, -lx, -ly};
    double lc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < 3; i++) {
        double L[4], R[4], dir[3];
        int kind;
        adj_edge_inputs(m, E, k, i, v[i], L, R, kind, dir);
        for (int pass = 0; pass < 3; pass++) {
            if (pass == 1 && kind >= 0) continue;   // no neighbour edge values
            for (int s0 = 0; s0 < 4; s0 += KF) {
                D Ld[4], Rd[4], Cd[4];
                for (int j = 0; j < 4; j++) {
                    Ld[j] = D(L[j]);
                    Rd[j] = D(R[j]);
                    Cd[j] = D(c[j]);
                }
                D *seed = pass == 0 ? Ld : (pass == 1 ? Rd : Cd);
                for (int j = 0; j < KF; j++) seed[s0 + j].d[j] = 1.0;
                D o[5];
                adj_edge_terms<D>(m, k, i, Ld, Rd, Cd, kind, dir, o);
                double gj[KF];
                for (int t = 0; t < KF; t++) {
                    gj[t] = 0.0;
                    for (int r = 0; r < 5; r++) gj[t] += lo[r] * o[r].d[t];
                }
                if (pass == 2) {
                    for (int t = 0; t < KF; t++) lc[s0 + t] += gj[t];
                } else {
                    const size_t slot = pass == 0 ? (size_t)i * N + k
                                                  : (size_t)(v[i] & 3) * N + (v[i] >> 2);
                    for (int t = 0; t < KF; t++)
                        if (gj[t] != 0.0) adj_add(lamE + (size_t)(s0 + t) * 3 * N + slot, gj[t]);
                }
            }
        }
    }
    adj_scatter_centroid(f, lc, lamS, zbar, N, k);
}

// ---- reverse-mode products of the extrapolation ----------------------------
// The same derivatives as the dual-number instantiation of adj_extrap_cell
// (the branches the forward takes; min / max pass the derivative of the
// argument they return, ties to the first; the limiter factor held where
// it is frozen), accumulated from the outputs back to the inputs in one
// sweep instead of one forward pass per group of input tangents.

// adj_limit's reverse: g (adjoint of the limited jumps, in) -> gd (of the
// unlimited jumps d), gqmin, gqmax, gbeta (added to).
ADJ_HD void adj_rev_limit(const double d[3], double qmin, double qmax, double beta,
                          double freeze, const double g[3], double gd[3], double &gqmin,
                          double &gqmax, double &gbeta)
{
    const double tiny = 1.0e-100;
    double r = 1000.0, r0 = 1.0;
    int rsrc = -1, r0src = -1, rnum = 0, r0num = 0;
    for (int i = 0; i < 3; i++) {
        if (d[i] < -tiny) {
            r0 = qmin / d[i];
            r0src = i;
            r0num = 1;
        }
        if (d[i] > tiny) {
            r0 = qmax / d[i];
            r0src = i;
            r0num = 2;
        }
        if (r0 <= r) {
            r = r0;
            rsrc = r0src;
            rnum = r0num;
        }
    }
    const double rb = r * beta;
    const bool live = rb <= 1.0 &&
                      !(freeze > 0.0 && fabs(d[0]) < freeze && fabs(d[1]) < freeze &&
                        fabs(d[2]) < freeze);
    const double phi = rb <= 1.0 ? rb : 1.0;
    double gphi = 0.0;
    for (int i = 0; i < 3; i++) {
        gd[i] = g[i] * phi;
        gphi += g[i] * d[i];
    }
    if (!live) return;
    const double gr = gphi * beta;
    gbeta += gphi * r;
    if (rsrc < 0) return;
    const double ib = 1.0 / d[rsrc];
    const double gnum = gr * ib;
    gd[rsrc] += -(r * gr) * ib;
    if (rnum == 1) gqmin += gnum;
    else gqmax += gnum;
}

// adj_aux_jumps' reverse: g (adjoint of d) -> gq[4] (qk, q0, q1, q2), gbeta
// (added to)
ADJ_HD void adj_rev_aux(double qk, double q0, double q1, double q2, double dx1, double dx2,
                        double dy1, double dy2, double inv_area2, const double dxv[3],
                        const double dyv[3], double beta, double freeze, const double g[3],
                        double gq[4], double &gbeta)
{
    const double dq0 = q0 - qk, dq1 = q1 - q0, dq2 = q2 - q0;
    double a = dy2 * dq1 - dy1 * dq2;
    a *= inv_area2;
    double b = dx1 * dq2 - dx2 * dq1;
    b *= inv_area2;
    double d[3];
    for (int i = 0; i < 3; i++) d[i] = a * dxv[i] + b * dyv[i];
    const double s1 = dq0 + dq1, s2 = dq0 + dq2;
    // qmax = max(max(dq0, max(s1, s2)), 0), each max passing its first on ties
    const int mx_in = s1 >= s2 ? 1 : 2;
    const double mxv = mx_in == 1 ? s1 : s2;
    const int mx_o = dq0 >= mxv ? 0 : mx_in;
    const double mxo = mx_o == 0 ? dq0 : mxv;
    const bool mx_live = mxo >= 0.0;
    const double qmax = mx_live ? mxo : 0.0;
    const int mn_in = s1 <= s2 ? 1 : 2;
    const double mnv = mn_in == 1 ? s1 : s2;
    const int mn_o = dq0 <= mnv ? 0 : mn_in;
    const double mno = mn_o == 0 ? dq0 : mnv;
    const bool mn_live = mno <= 0.0;
    const double qmin = mn_live ? mno : 0.0;
    double gd[3], gqmin = 0.0, gqmax = 0.0;
    adj_rev_limit(d, qmin, qmax, beta, freeze, g, gd, gqmin, gqmax, gbeta);
    double gdq0 = 0.0, gs1 = 0.0, gs2 = 0.0;
    if (mx_live) {
        if (mx_o == 0) gdq0 += gqmax;
        else if (mx_o == 1) gs1 += gqmax;
        else gs2 += gqmax;
    }
    if (mn_live) {
        if (mn_o == 0) gdq0 += gqmin;
        else if (mn_o == 1) gs1 += gqmin;
        else gs2 += gqmin;
    }
    double ga = 0.0, gb = 0.0;
    for (int i = 0; i < 3; i++) {
        ga += gd[i] * dxv[i];
        gb += gd[i] * dyv[i];
    }
    ga *= inv_area2;
    gb *= inv_area2;
    const double gdq1 = ga * dy2 - gb * dx2 + gs1;
    const double gdq2 = gb * dx1 - ga * dy1 + gs2;
    gdq0 += gs1 + gs2;
    gq[0] -= gdq0;
    gq[1] += gdq0 - gdq1 - gdq2;
    gq[2] += gdq1;
    gq[3] += gdq2;
}

// adj_line_jumps' reverse: g (adjoint of d) -> gqk, gq1 (added to); beta is a
// constant there
ADJ_HD void adj_rev_line(double qk, double q1, double dx2, double dy2, const double dxv[3],
                         const double dyv[3], double beta, double freeze, const double g[3],
                         double &gqk, double &gq1)
{
    const double dq1 = q1 - qk;
    const double a = dq1 * dx2, b = dq1 * dy2;
    double d[3];
    for (int i = 0; i < 3; i++) d[i] = a * dxv[i] + b * dyv[i];
    const bool up = dq1 >= 0.0;
    const double qmin = up ? 0.0 : dq1, qmax = up ? dq1 : 0.0;
    double gd[3], gqmin = 0.0, gqmax = 0.0, gbeta = 0.0;
    adj_rev_limit(d, qmin, qmax, beta, freeze, g, gd, gqmin, gqmax, gbeta);
    double gdq1 = up ? gqmax : gqmin;
    for (int i = 0; i < 3; i++) gdq1 += (gd[i] * dxv[i]) * dx2 + (gd[i] * dyv[i]) * dy2;
    gq1 += gdq1;
    gqk -= gdq1;
}

// The vector-Jacobian product of adj_extrap_cell: lo[4 i + t] (adjoints of
// ew, eh, euh, evh on edge i) -> gs[4 j + c] (adjoints of w, uh, vh, z of
// stencil slot j).
ADJ_HD void adj_extrap_vjp(const double w[4], const double uh[4], const double vh[4],
                           const double z[4], const unsigned fl[4], const AdjExtrapGeo &gm,
                           const AdjParams &p, const double lo[12], double gs[16])
{
    for (int j = 0; j < 16; j++) gs[j] = 0.0;
    // forward values: the cells
    const bool clamp0 = fl[0] & ADJ_F_CLAMP, keep0 = fl[0] & ADJ_F_KEEP;
    const double cw = clamp0 ? z[0] : w[0];
    const bool ch_live = cw - z[0] >= 0.0;
    const double ch = ch_live ? cw - z[0] : 0.0;
    double cu = 0.0, cv = 0.0, cinv = 0.0;
    if (keep0) {
        cinv = 1.0 / ch;
        cu = uh[0] * cinv;
        cv = vh[0] * cinv;
    }
    double nw[3], nh[3], nu[3], nv[3], ninv[3];
    bool nh_live[3];
    for (int i = 0; i < 3; i++) {
        nh_live[i] = w[i + 1] - z[i + 1] >= 0.0;
        nh[i] = nh_live[i] ? w[i + 1] - z[i + 1] : 0.0;
        nw[i] = (fl[i + 1] & ADJ_F_CLAMP) ? z[i + 1] : w[i + 1];
        nu[i] = nv[i] = ninv[i] = 0.0;
        if (fl[i + 1] & ADJ_F_KEEP) {
            ninv[i] = 1.0 / nh[i];
            nu[i] = uh[i + 1] * ninv[i];
            nv[i] = vh[i + 1] * ninv[i];
        }
    }
    // forward values: the jumps (limited), as adj_extrap_cell forms them
    double jw[3] = {0.0, 0.0, 0.0}, jh[3] = {0.0, 0.0, 0.0};
    double ju[3] = {0.0, 0.0, 0.0}, jv[3] = {0.0, 0.0, 0.0};
    bool constant = gm.nbound == 3, gw = true, gu = true, gv = true;
    bool aux = false, line = false;
    double area2 = 0.0, inv_area2 = 0.0, hfactor = 0.0, bw_ = 0.0, bu_ = 0.0, bv_ = 0.0;
    double ldx2 = 0.0, ldy2 = 0.0;
    int lsel = 0;
    const double c_tmp = 1.0 / (0.3 - 0.1), d_tmp = 1.0 - (c_tmp * 0.3);
    // hfactor's selections
    int hmin_src = 0, hmax_src = 0;
    double hmin = 0.0, hmax = 0.0;
    bool t1_live = false, t2_live = false, t3_live = false, t4_live = false, m1_live = false;
    bool m2_a = false, hf1_live = false, t5_live = false, t6_live = false, c_sel = false;
    double t1 = 0.0, t2 = 0.0, t3 = 0.0, t4 = 0.0, rA = 0.0, rB = 0.0, m1 = 0.0, m2 = 0.0;
    double t5 = 0.0, t6 = 0.0, C = 0.0, hf1 = 0.0;
    if (gm.nbound <= 1) {
        area2 = gm.dy2 * gm.dx1 - gm.dy1 * gm.dx2;
        if (area2 <= 0.0) {
            constant = true;
        } else {
            aux = true;
            // hmin = min(min(nh0, min(nh1, nh2)), ch); hmax likewise
            {
                const int ia = nh[1] <= nh[2] ? 1 : 2;
                const int ib = nh[0] <= nh[ia] ? 0 : ia;
                hmin_src = nh[ib] <= ch ? ib : 3;
                hmin = hmin_src == 3 ? ch : nh[ib];
                const int ja = nh[1] >= nh[2] ? 1 : 2;
                const int jb = nh[0] >= nh[ja] ? 0 : ja;
                hmax_src = nh[jb] >= ch ? jb : 3;
                hmax = hmax_src == 3 ? ch : nh[jb];
            }
            t1_live = hmin >= 0.0;
            t1 = t1_live ? hmin : 0.0;
            t2_live = ch >= 1.0e-06;
            t2 = t2_live ? ch : 1.0e-06;
            rA = c_tmp * t1 / t2;
            const double A = rA + d_tmp;
            t3_live = ch >= 0.0;
            t3 = t3_live ? ch : 0.0;
            t4_live = hmax >= 1.0e-06;
            t4 = t4_live ? hmax : 1.0e-06;
            rB = c_tmp * t3 / t4;
            const double B = rB + d_tmp;
            m1_live = B <= 1.0;
            m1 = m1_live ? B : 1.0;
            m2_a = A <= m1;
            m2 = m2_a ? A : m1;
            hf1_live = !(0.0 >= m2);
            hf1 = hf1_live ? m2 : 0.0;
            t5_live = hmin - p.mah >= 0.0;
            t5 = t5_live ? hmin - p.mah : 0.0;
            t6_live = hmin >= 0.0;
            t6 = (t6_live ? hmin : 0.0) + 1.0 * p.mah;
            C = 1.2 * t5 / t6;
            c_sel = C <= hf1;
            hfactor = c_sel ? C : hf1;
            inv_area2 = 1.0 / area2;
            bw_ = p.bwd + (p.bw - p.bwd) * hfactor;
            gw = bw_ > 0.0;
            if (gw) {
                adj_aux_jumps<double>(cw, nw[0], nw[1], nw[2], gm.dx1, gm.dx2, gm.dy1, gm.dy2,
                                      inv_area2, gm.dxv, gm.dyv, bw_, jw, p.freeze);
                adj_aux_jumps<double>(ch, nh[0], nh[1], nh[2], gm.dx1, gm.dx2, gm.dy1, gm.dy2,
                                      inv_area2, gm.dxv, gm.dyv, bw_, jh, p.freeze);
            }
            bu_ = p.buhd + (p.buh - p.buhd) * hfactor;
            gu = bu_ > 0.0;
            if (gu)
                adj_aux_jumps<double>(cu, nu[0], nu[1], nu[2], gm.dx1, gm.dx2, gm.dy1, gm.dy2,
                                      inv_area2, gm.dxv, gm.dyv, bu_, ju, p.freeze);
            bv_ = p.bvhd + (p.bvh - p.bvhd) * hfactor;
            gv = bv_ > 0.0;
            if (gv)
                adj_aux_jumps<double>(cv, nv[0], nv[1], nv[2], gm.dx1, gm.dx2, gm.dy1, gm.dy2,
                                      inv_area2, gm.dxv, gm.dyv, bv_, jv, p.freeze);
        }
    } else if (gm.nbound == 2) {
        if (!(gm.interior[0] || gm.interior[1] || gm.interior[2])) {
            constant = true;
        } else {
            line = true;
            lsel = gm.interior[0] ? 0 : (gm.interior[1] ? 1 : 2);
            const double la2 = gm.lx * gm.lx + gm.ly * gm.ly;
            ldx2 = 1.0 / la2;
            ldy2 = ldx2 * gm.ly;
            ldx2 *= gm.lx;
            adj_line_jumps<double>(cw, nw[lsel], ldx2, ldy2, gm.dxv, gm.dyv, p.bw, jw, p.freeze);
            adj_line_jumps<double>(ch, nh[lsel], ldx2, ldy2, gm.dxv, gm.dyv, p.bw, jh, p.freeze);
            adj_line_jumps<double>(cu, nu[lsel], ldx2, ldy2, gm.dxv, gm.dyv, p.bw, ju, p.freeze);
            adj_line_jumps<double>(cv, nv[lsel], ldx2, ldy2, gm.dxv, gm.dyv, p.bw, jv, p.freeze);
        }
    }
    // reverse: the edge values
    const bool kw = !(constant || !gw), ku = !(constant || !gu), kv = !(constant || !gv);
    double gcw = 0.0, gch = 0.0, gcu = 0.0, gcv = 0.0;
    double gjw[3], gjh[3], gju[3], gjv[3];
    for (int i = 0; i < 3; i++) {
        const double eh = kw ? ch + jh[i] : ch;
        const double eu = ku ? cu + ju[i] : cu;
        const double ev = kv ? cv + jv[i] : cv;
        const double lw = lo[4 * i], lh = lo[4 * i + 1], lu = lo[4 * i + 2], lv = lo[4 * i + 3];
        const double geh = lh + lu * eu + lv * ev;
        const double geu = lu * eh, gev = lv * eh;
        gcw += lw;
        gch += geh;
        gcu += geu;
        gcv += gev;
        gjw[i] = kw ? lw : 0.0;
        gjh[i] = kw ? geh : 0.0;
        gju[i] = ku ? geu : 0.0;
        gjv[i] = kv ? gev : 0.0;
    }
    // reverse: the jumps
    double gnw[3] = {0.0, 0.0, 0.0}, gnh[3] = {0.0, 0.0, 0.0};
    double gnu[3] = {0.0, 0.0, 0.0}, gnv[3] = {0.0, 0.0, 0.0};
    if (aux) {
        double ghf = 0.0;
        if (gw) {
            double gq[4] = {0.0, 0.0, 0.0, 0.0}, gb = 0.0;
            adj_rev_aux(cw, nw[0], nw[1], nw[2], gm.dx1, gm.dx2, gm.dy1, gm.dy2, inv_area2,
                        gm.dxv, gm.dyv, bw_, p.freeze, gjw, gq, gb);
            gcw += gq[0];
            for (int i = 0; i < 3; i++) gnw[i] += gq[i + 1];
            double gq2[4] = {0.0, 0.0, 0.0, 0.0};
            adj_rev_aux(ch, nh[0], nh[1], nh[2], gm.dx1, gm.dx2, gm.dy1, gm.dy2, inv_area2,
                        gm.dxv, gm.dyv, bw_, p.freeze, gjh, gq2, gb);
            gch += gq2[0];
            for (int i = 0; i < 3; i++) gnh[i] += gq2[i + 1];
            ghf += gb * (p.bw - p.bwd);
        }
        if (gu) {
            double gq[4] = {0.0, 0.0, 0.0, 0.0}, gb = 0.0;
            adj_rev_aux(cu, nu[0], nu[1], nu[2], gm.dx1, gm.dx2, gm.dy1, gm.dy2, inv_area2,
                        gm.dxv, gm.dyv, bu_, p.freeze, gju, gq, gb);
            gcu += gq[0];
            for (int i = 0; i < 3; i++) gnu[i] += gq[i + 1];
            ghf += gb * (p.buh - p.buhd);
        }
        if (gv) {
            double gq[4] = {0.0, 0.0, 0.0, 0.0}, gb = 0.0;
            adj_rev_aux(cv, nv[0], nv[1], nv[2], gm.dx1, gm.dx2, gm.dy1, gm.dy2, inv_area2,
                        gm.dxv, gm.dyv, bv_, p.freeze, gjv, gq, gb);
            gcv += gq[0];
            for (int i = 0; i < 3; i++) gnv[i] += gq[i + 1];
            ghf += gb * (p.bvh - p.bvhd);
        }
        // hfactor -> ch, hmin, hmax
        double ghmin = 0.0, ghmax = 0.0;
        const double gC = c_sel ? ghf : 0.0, ghf1 = c_sel ? 0.0 : ghf;
        // C = 1.2 t5 / t6
        const double gt5 = gC * 1.2 / t6;
        const double gt6 = -(gC * C) / t6;
        if (t5_live) ghmin += gt5;
        if (t6_live) ghmin += gt6;
        const double gm2 = hf1_live ? ghf1 : 0.0;
        const double gA = m2_a ? gm2 : 0.0, gm1 = m2_a ? 0.0 : gm2;
        const double gB = m1_live ? gm1 : 0.0;
        // A = c t1 / t2 + d, B = c t3 / t4 + d
        const double gt1 = gA * c_tmp / t2, gt2 = -(gA * rA) / t2;
        const double gt3 = gB * c_tmp / t4, gt4 = -(gB * rB) / t4;
        if (t1_live) ghmin += gt1;
        if (t2_live) gch += gt2;
        if (t3_live) gch += gt3;
        if (t4_live) ghmax += gt4;
        if (hmin_src == 3) gch += ghmin;
        else gnh[hmin_src] += ghmin;
        if (hmax_src == 3) gch += ghmax;
        else gnh[hmax_src] += ghmax;
    } else if (line) {
        double gqk = 0.0, gq1 = 0.0;
        adj_rev_line(cw, nw[lsel], ldx2, ldy2, gm.dxv, gm.dyv, p.bw, p.freeze, gjw, gqk, gq1);
        gcw += gqk;
        gnw[lsel] += gq1;
        gqk = gq1 = 0.0;
        adj_rev_line(ch, nh[lsel], ldx2, ldy2, gm.dxv, gm.dyv, p.bw, p.freeze, gjh, gqk, gq1);
        gch += gqk;
        gnh[lsel] += gq1;
        gqk = gq1 = 0.0;
        adj_rev_line(cu, nu[lsel], ldx2, ldy2, gm.dxv, gm.dyv, p.bw, p.freeze, gju, gqk, gq1);
        gcu += gqk;
        gnu[lsel] += gq1;
        gqk = gq1 = 0.0;
        adj_rev_line(cv, nv[lsel], ldx2, ldy2, gm.dxv, gm.dyv, p.bw, p.freeze, gjv, gqk, gq1);
        gcv += gqk;
        gnv[lsel] += gq1;
    }
    // reverse: the cells -> the raw state of each slot
    if (keep0) {
        gs[1] += gcu * cinv;
        gs[2] += gcv * cinv;
        const double ginv = gcu * uh[0] + gcv * vh[0];
        gch += -(ginv * cinv) * cinv;
    }
    if (ch_live) {
        gcw += gch;
        gs[3] -= gch;
    }
    if (clamp0) gs[3] += gcw;
    else gs[0] += gcw;
    for (int i = 0; i < 3; i++) {
        const int j = 4 * (i + 1);
        double gh = gnh[i];
        if (fl[i + 1] & ADJ_F_KEEP) {
            gs[j + 1] += gnu[i] * ninv[i];
            gs[j + 2] += gnv[i] * ninv[i];
            const double ginv = gnu[i] * uh[i + 1] + gnv[i] * vh[i + 1];
            gh += -(ginv * ninv[i]) * ninv[i];
        }
        if (nh_live[i]) {
            gs[j] += gh;
            gs[j + 3] -= gh;
        }
        if (fl[i + 1] & ADJ_F_CLAMP) gs[j + 3] += gnw[i];
        else gs[j] += gnw[i];
    }
}

// ---- reverse-mode product of one edge's update terms -------------------------
// adj_flux_central's reverse: gef (adjoints of the flux), gpf (of the pressure
// flux) -> adjoints of its inputs, added to g[11] in the order ql0, ql1, ql2,
// qr0, qr1, qr2, h_left, h_right, hle, hre, ze.
ADJ_HD void adj_rev_flux_central(double ql0, double ql1, double ql2, double qr0, double qr1,
                                 double qr2, double h_left, double h_right, double hle,
                                 double hre, double n1, double n2, double eps, double ze,
                                 double g, const double gef[3], double gpf, double gq[11])
{
    if (h_left == 0.0 && h_right == 0.0) return;
    const double l1 = n1 * ql1 + n2 * ql2, l2 = (-n2) * ql1 + n1 * ql2;
    const double r1 = n1 * qr1 + n2 * qr2, r2 = (-n2) * qr1 + n1 * qr2;
    const bool lp = hle > 0.0, rp = hre > 0.0;
    double tl = 0.0, ul = 0.0, uhl = 0.0, vhl = 0.0, tr = 0.0, ur = 0.0, uhr = 0.0, vhr = 0.0;
    if (lp) {
        tl = 1.0 / hle;
        ul = l1 * tl;
        uhl = h_left * ul;
        vhl = h_left * tl * l2;
    }
    if (rp) {
        tr = 1.0 / hre;
        ur = r1 * tr;
        uhr = h_right * ur;
        vhr = h_right * tr * r2;
    }
    const double cl = sqrt(g * h_left), cr = sqrt(g * h_right);
    const double a1 = ul + cl, a2 = ur + cr, b1 = ul - cl, b2 = ur - cr;
    const bool smax_l = a1 >= a2, smin_l = b1 <= b2;
    double s_max = smax_l ? a1 : a2, s_min = smin_l ? b1 : b2;
    const bool smax_live = !(s_max < 0.0), smin_live = !(s_min > 0.0);
    if (!smax_live) s_max = 0.0;
    if (!smin_live) s_min = 0.0;
    double gul = 0.0, guhl = 0.0, gvhl = 0.0, gur = 0.0, guhr = 0.0, gvhr = 0.0;
    double gsmax = 0.0, gsmin = 0.0, ghl = 0.0, ghr = 0.0;
    const double denom = s_max - s_min;
    if (denom < eps) {
        // ef = 0; pflux = 0.5 g 0.5 (hl^2 + hr^2)
        const double c = 0.5 * g * 0.5;
        gq[6] += gpf * c * (h_left + h_left);
        gq[7] += gpf * c * (h_right + h_right);
        return;
    }
    const double fl0 = ul * h_left, fl1 = ul * uhl, fl2 = ul * vhl;
    const double fr0 = ur * h_right, fr1 = ur * uhr, fr2 = ur * vhr;
    const double inv = 1.0 / denom;
    const double sms = s_max * s_min;
    const bool mr_q = qr0 >= ze, ml_q = ql0 >= ze;
    const double mr = mr_q ? qr0 : ze, ml = ml_q ? ql0 : ze;
    const double E0 = s_max * fl0 - s_min * fr0 + sms * (mr - ml);
    const double E1 = s_max * fl1 - s_min * fr1 + 1.0 * sms * (uhr - uhl);
    const double E2 = s_max * fl2 - s_min * fr2 + 1.0 * sms * (vhr - vhl);
    const double Q = s_max * h_left * h_left - s_min * h_right * h_right;
    const double P = 0.5 * g * Q;
    // ef0 = e0, ef1 = n1 e1 - n2 e2, ef2 = n2 e1 + n1 e2
    const double ge0 = gef[0], ge1 = n1 * gef[1] + n2 * gef[2];
    const double ge2 = (-n2) * gef[1] + n1 * gef[2];
    const double gE0 = ge0 * inv, gE1 = ge1 * inv, gE2 = ge2 * inv, gP = gpf * inv;
    const double ginv = ge0 * E0 + ge1 * E1 + ge2 * E2 + gpf * P;
    const double gden = -(ginv * inv) * inv;
    gsmax += gden;
    gsmin -= gden;
    double gsms = 0.0;
    gsmax += gE0 * fl0 + gE1 * fl1 + gE2 * fl2;
    gsmin -= gE0 * fr0 + gE1 * fr1 + gE2 * fr2;
    const double gfl0 = gE0 * s_max, gfl1 = gE1 * s_max, gfl2 = gE2 * s_max;
    const double gfr0 = -gE0 * s_min, gfr1 = -gE1 * s_min, gfr2 = -gE2 * s_min;
    gsms += gE0 * (mr - ml) + gE1 * (uhr - uhl) + gE2 * (vhr - vhl);
    const double gm = gE0 * sms;
    if (mr_q) gq[3] += gm;
    else gq[10] += gm;
    if (ml_q) gq[0] -= gm;
    else gq[10] -= gm;
    guhr += gE1 * sms;
    guhl -= gE1 * sms;
    gvhr += gE2 * sms;
    gvhl -= gE2 * sms;
    const double gQ = gP * (0.5 * g);
    gsmax += gQ * (h_left * h_left);
    gsmin -= gQ * (h_right * h_right);
    ghl += gQ * (s_max * h_left + s_max * h_left);
    ghr -= gQ * (s_min * h_right + s_min * h_right);
    gsmax += gsms * s_min;
    gsmin += gsms * s_max;
    gul += gfl0 * h_left + gfl1 * uhl + gfl2 * vhl;
    ghl += gfl0 * ul;
    guhl += gfl1 * ul;
    gvhl += gfl2 * ul;
    gur += gfr0 * h_right + gfr1 * uhr + gfr2 * vhr;
    ghr += gfr0 * ur;
    guhr += gfr1 * ur;
    gvhr += gfr2 * ur;
    double gcl = 0.0, gcr = 0.0;
    if (smax_live) {
        if (smax_l) {
            gul += gsmax;
            gcl += gsmax;
        } else {
            gur += gsmax;
            gcr += gsmax;
        }
    }
    if (smin_live) {
        if (smin_l) {
            gul += gsmin;
            gcl -= gsmin;
        } else {
            gur += gsmin;
            gcr -= gsmin;
        }
    }
    if (cl > 0.0) ghl += gcl * (0.5 / cl) * g;
    if (cr > 0.0) ghr += gcr * (0.5 / cr) * g;
    double gl1 = 0.0, gl2 = 0.0, gr1 = 0.0, gr2 = 0.0;
    if (lp) {
        const double ght = gvhl * l2;
        gl2 += gvhl * (h_left * tl);
        ghl += ght * tl;
        double gtl = ght * h_left;
        ghl += guhl * ul;
        gul += guhl * h_left;
        gl1 += gul * tl;
        gtl += gul * l1;
        gq[8] += -(gtl * tl) / hle;
    }
    if (rp) {
        const double ght = gvhr * r2;
        gr2 += gvhr * (h_right * tr);
        ghr += ght * tr;
        double gtr = ght * h_right;
        ghr += guhr * ur;
        gur += guhr * h_right;
        gr1 += gur * tr;
        gtr += gur * r1;
        gq[9] += -(gtr * tr) / hre;
    }
    gq[1] += gl1 * n1 + gl2 * (-n2);
    gq[2] += gl1 * n2 + gl2 * n1;
    gq[4] += gr1 * n1 + gr2 * (-n2);
    gq[5] += gr1 * n2 + gr2 * n1;
    gq[6] += ghl;
    gq[7] += ghr;
}

// The vector-Jacobian product of adj_edge_terms: go[5] (adjoints of the five
// update terms) -> gL, gR, gC (adjoints of the own and neighbour edge values
// w, h, uh, vh and of the centroid w, uh, vh, z), added to. The same
// derivatives as the dual-number instantiation.
ADJ_HD void adj_edge_vjp(const AdjView &m, int k, int i, const double L[4], const double R[4],
                         const double C[4], int kind, const double dir[3], const double go[5],
                         double gL[4], double gR[4], double gC[4])
{
    const size_t N = m.N, ki = (size_t)i * N + k;
    const double n1 = m.nx[ki], n2 = m.ny[ki], len = m.len[ki];
    const AdjParams &p = m.p;
    const bool hc_live = C[0] - C[3] >= 0.0;
    const double hc = hc_live ? C[0] - C[3] : 0.0, zc = C[3];
    const double ql0 = L[0], hle = L[1], ql1 = L[2], ql2 = L[3];
    const double zl = ql0 - hle;
    double qr0, qr1, qr2, zr, hre;
    double r1 = 0.0, r2 = 0.0;
    bool hre_live = true;
    if (kind < 0) {
        qr0 = R[0];
        qr1 = R[2];
        qr2 = R[3];
        hre = R[1];
        zr = qr0 - hre;
    } else {
        if (kind == 0) {
            r1 = -ql1 * n1 - ql2 * n2;
            r2 = -ql1 * n2 + ql2 * n1;
            qr0 = ql0;
            qr1 = n1 * r1 - n2 * r2;
            qr2 = n2 * r1 + n1 * r2;
        } else if (kind == 1) {
            qr0 = ql0;
            qr1 = ql1;
            qr2 = ql2;
        } else if (kind == 2) {
            qr0 = dir[0];
            qr1 = dir[1];
            qr2 = dir[2];
        } else {
            qr0 = C[0];
            qr1 = C[1];
            qr2 = C[2];
        }
        zr = zl;
        hre_live = qr0 - zr >= 0.0;
        hre = hre_live ? qr0 - zr : 0.0;
    }
    const bool zh_l = zl >= zr;
    const double z_half = zh_l ? zl : zr;
    const double hl_in = hle + zl - z_half, hr_in = hre + zr - z_half;
    const bool hl_live = hl_in >= 0.0, hr_live = hr_in >= 0.0;
    const double h_left = hl_live ? hl_in : 0.0, h_right = hr_live ? hr_in : 0.0;
    // forward pressure flux (needed by nothing below but its adjoint's path)
    // out = -ef len (3), n1 pg, n2 pg; pg = len (K X + pflux), K = -g 0.5
    const double gpg = go[3] * n1 + go[4] * n2;
    const double gef[3] = {-go[0] * len, -go[1] * len, -go[2] * len};
    const double K = -p.g * 0.5;
    const double gX = (gpg * len) * K, gpf = gpg * len;
    double gq[11] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    adj_rev_flux_central(ql0, ql1, ql2, qr0, qr1, qr2, h_left, h_right, hle, hre, n1, n2, p.eps,
                         z_half, p.g, gef, gpf, gq);
    // X = h_left^2 - hle^2 - (hle + hc)(zl - zc)
    double ghl = gq[6] + gX * (h_left + h_left);
    double ghle = gq[8] - gX * (hle + hle) - gX * (zl - zc);
    const double ghc = -gX * (zl - zc);
    double gzl = -gX * (hle + hc);
    const double gzc = gX * (hle + hc);
    double ghr = gq[7], ghre = gq[9], gzh = gq[10], gzr = 0.0;
    double gql0 = gq[0], gql1 = gq[1], gql2 = gq[2], gqr0 = gq[3], gqr1 = gq[4], gqr2 = gq[5];
    if (hr_live) {
        ghre += ghr;
        gzr += ghr;
        gzh -= ghr;
    }
    if (hl_live) {
        ghle += ghl;
        gzl += ghl;
        gzh -= ghl;
    }
    if (zh_l) gzl += gzh;
    else gzr += gzh;
    if (kind < 0) {
        gR[0] += gqr0 + gzr;
        gR[1] += ghre - gzr;
        gR[2] += gqr1;
        gR[3] += gqr2;
    } else {
        if (hre_live) {
            gqr0 += ghre;
            gzr -= ghre;
        }
        gzl += gzr;
        if (kind == 0) {
            gql0 += gqr0;
            const double gr1 = gqr1 * n1 + gqr2 * n2;
            const double gr2 = -gqr1 * n2 + gqr2 * n1;
            gql1 += -gr1 * n1 - gr2 * n2;
            gql2 += -gr1 * n2 + gr2 * n1;
        } else if (kind == 1) {
            gql0 += gqr0;
            gql1 += gqr1;
            gql2 += gqr2;
        } else if (kind != 2) {
            gC[0] += gqr0;
            gC[1] += gqr1;
            gC[2] += gqr2;
        }
    }
    gql0 += gzl;
    ghle -= gzl;
    gL[0] += gql0;
    gL[1] += ghle;
    gL[2] += gql1;
    gL[3] += gql2;
    if (hc_live) {
        gC[0] += ghc;
        gC[3] -= ghc;
    }
    gC[3] += gzc;
}

// adj_bwd_flux by reverse mode: the same additions from one sweep of
// adj_edge_vjp per edge.
ADJ_HD void adj_bwd_flux_rev(const AdjView &m, const double *S, const unsigned char *fl,
                             const double *E, const double *lam_ex, int k, double *lamE,
                             double *lamS, double *zbar)
{
    const int N = m.N;
    int v[3], sur[3];
    adj_neighbours(m, k, v, sur);
    double q[4] = {S[4 * k], S[4 * k + 1], S[4 * k + 2], S[4 * k + 3]}, c[4];
    const unsigned f = fl[k];
    adj_centroid<double>(q, f, c);
    const double inv_area = 1.0 / m.area[k];
    const double lw = lam_ex[k] * inv_area, lx = lam_ex[N + k] * inv_area;
    const double ly = lam_ex[2 * (size_t)N + k] * inv_area;
    const double lo[5] = {lw, lx, ly, -lx, -ly};
    double lc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < 3; i++) {
        double L[4], R[4], dir[3];
        int kind;
        adj_edge_inputs(m, E, k, i, v[i], L, R, kind, dir);
        double gL[4] = {0.0, 0.0, 0.0, 0.0}, gR[4] = {0.0, 0.0, 0.0, 0.0};
        double gC[4] = {0.0, 0.0, 0.0, 0.0};
        adj_edge_vjp(m, k, i, L, R, c, kind, dir, lo, gL, gR, gC);
        const size_t own = (size_t)i * N + k;
        for (int t = 0; t < 4; t++)
            if (gL[t] != 0.0) adj_add(lamE + (size_t)t * 3 * N + own, gL[t]);
        if (kind < 0) {
            const size_t nbs = (size_t)(v[i] & 3) * N + (v[i] >> 2);
            for (int t = 0; t < 4; t++)
                if (gR[t] != 0.0) adj_add(lamE + (size_t)t * 3 * N + nbs, gR[t]);
        }
        for (int t = 0; t < 4; t++) lc[t] += gC[t];
    }
    adj_scatter_centroid(f, lc, lamS, zbar, N, k);
}

// Extrapolation of triangle k: lamE (its edge values) -> the adjoints of the
// raw state of k and of its surrogate neighbours (and of their beds), stored
// per stencil slot: X[(4 slot + c) N + k], slot 0 for k itself (a boundary
// edge's surrogate is k and counts there), slot 1 + i for the neighbour
// across edge i, c over w, uh, vh, z. adj_gather_extrap then sums each
// triangle's contributions in a fixed order, so the result does not depend on
// the order in which triangles run. K tangents per pass (the four components
// of one stencil triangle in 4 / K passes): a narrower dual keeps the long
// extrapolation in registers.
template <int K = 4>
ADJ_HD void adj_bwd_extrap(const AdjView &m, const double *S, const unsigned char *fl,
                           const double *lamE, int k, double *X)
{
    typedef Dual<K> D;
    const int N = m.N;
    int v[3], sur[3];
    adj_neighbours(m, k, v, sur);
    const AdjExtrapGeo g = adj_geometry(m, k, v, sur);
    const int id[4] = {k, sur[0], sur[1], sur[2]};
    double lo[12];
    bool any = false;
    for (int i = 0; i < 3; i++)
        for (int t = 0; t < 4; t++) {
            lo[4 * i + t] = lamE[(size_t)t * 3 * N + (size_t)i * N + k];
            any = any || lo[4 * i + t] != 0.0;
        }
    for (int j = 0; j < 16; j++) X[(size_t)j * N + k] = 0.0;
    if (!any) return;
    unsigned f[4];
    for (int j = 0; j < 4; j++) f[j] = fl[id[j]];
    // passes over the distinct triangles of the stencil (k, then each interior
    // neighbour) and over groups of K of their four components
#pragma unroll 1
    for (int pass = 0; pass < 4 * (4 / K); pass++) {
        const int slot = pass / (4 / K), c0 = (pass % (4 / K)) * K;
        const int t = slot == 0 ? k : (slot == 1 ? sur[0] : (slot == 2 ? sur[1] : sur[2]));
        const int vs = slot == 1 ? v[0] : (slot == 2 ? v[1] : v[2]);
        if (slot > 0 && vs < 0) continue;   // a boundary surrogate is k itself
        // seeds by comparison, so that every array index is a constant after
        // unrolling and the duals stay in registers
        D w[4], uh[4], vh[4], z[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            w[j] = D(S[4 * id[j]]);
            uh[j] = D(S[4 * id[j] + 1]);
            vh[j] = D(S[4 * id[j] + 2]);
            z[j] = D(S[4 * id[j] + 3]);
            const bool mine = id[j] == t;
#pragma unroll
            for (int c = 0; c < K; c++) {
                const int cc = c0 + c;
                w[j].d[c] = (mine && cc == 0) ? 1.0 : 0.0;
                uh[j].d[c] = (mine && cc == 1) ? 1.0 : 0.0;
                vh[j].d[c] = (mine && cc == 2) ? 1.0 : 0.0;
                z[j].d[c] = (mine && cc == 3) ? 1.0 : 0.0;
            }
        }
        D ew[3], eh[3], euh[3], evh[3];
        adj_extrap_cell<D>(w, uh, vh, z, f, g, m.p, ew, eh, euh, evh);
#pragma unroll
        for (int c = 0; c < K; c++) {
            double a = 0.0;
#pragma unroll
            for (int i = 0; i < 3; i++)
                a += lo[4 * i] * ew[i].d[c] + lo[4 * i + 1] * eh[i].d[c] +
                     lo[4 * i + 2] * euh[i].d[c] + lo[4 * i + 3] * evh[i].d[c];
            X[(size_t)(4 * slot + c0 + c) * N + k] = a;
        }
    }
}

// adj_bwd_extrap by reverse mode: the same stores into X from one sweep of
// adj_extrap_vjp.
ADJ_HD void adj_bwd_extrap_rev(const AdjView &m, const double *S, const unsigned char *fl,
                               const double *lamE, int k, double *X)
{
    const int N = m.N;
    int v[3], sur[3];
    adj_neighbours(m, k, v, sur);
    double lo[12];
    bool any = false;
    for (int i = 0; i < 3; i++)
        for (int t = 0; t < 4; t++) {
            lo[4 * i + t] = lamE[(size_t)t * 3 * N + (size_t)i * N + k];
            any = any || lo[4 * i + t] != 0.0;
        }
    double out[16];
    for (int j = 0; j < 16; j++) out[j] = 0.0;
    if (any) {
        const AdjExtrapGeo g = adj_geometry(m, k, v, sur);
        const int id[4] = {k, sur[0], sur[1], sur[2]};
        double w[4], uh[4], vh[4], z[4];
        unsigned f[4];
        for (int j = 0; j < 4; j++) {
            w[j] = S[4 * id[j]];
            uh[j] = S[4 * id[j] + 1];
            vh[j] = S[4 * id[j] + 2];
            z[j] = S[4 * id[j] + 3];
            f[j] = fl[id[j]];
        }
        double gs[16];
        adj_extrap_vjp(w, uh, vh, z, f, g, m.p, lo, gs);
        // a boundary edge's surrogate is k itself: its slot's share is k's
        for (int c = 0; c < 4; c++) out[c] = gs[c];
        for (int i = 0; i < 3; i++)
            for (int c = 0; c < 4; c++) {
                if (v[i] >= 0) out[4 * (i + 1) + c] = gs[4 * (i + 1) + c];
                else out[c] += gs[4 * (i + 1) + c];
            }
    }
    for (int j = 0; j < 16; j++) X[(size_t)j * N + k] = out[j];
}

// The extrapolation adjoints that reach triangle t (adj_bwd_extrap's X): its
// own stencil's, then those of its neighbours across edges 0, 1, 2, added to
// the state adjoint lamS and the bed adjoint zbar.
ADJ_HD void adj_gather_extrap(const AdjView &m, const double *X, int t, double *lamS,
                              double *zbar)
{
    const int N = m.N;
    int v[3], sur[3];
    adj_neighbours(m, t, v, sur);
    for (int c = 0; c < 4; c++) {
        double a = X[(size_t)c * N + t];
        for (int i = 0; i < 3; i++)
            if (v[i] >= 0) a += X[(size_t)(4 * (1 + (v[i] & 3)) + c) * N + (v[i] >> 2)];
        if (a == 0.0) continue;
        if (c < 3) lamS[(size_t)c * N + t] += a;
        else zbar[t] += a;
    }
}

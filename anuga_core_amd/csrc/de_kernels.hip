// HIP kernels of the FP64 DE step for gfx950.
//
// One thread per triangle, triangles in locality (RCM) order, 256-thread
// workgroups remapped XCD-contiguous. Per Euler substep:
//   de_extrapolate  protect + velocity + dry-neighbour zeroing + limited
//                   second-order edge values (sw_domain_orig.c:1215-1281,
//                   :1994-2680). Reads the raw centroid state and never
//                   writes it: the protect / zeroing outcome goes to a flag
//                   byte, so the step's initial state doubles as the RK2 backup.
//   de_flux         per-triangle gather of its three edge fluxes in local edge
//                   order (bit-identical to the reference's compute-once-and-
//                   mirror loop, sw_domain_orig.c:765-1150), boundary states
//                   evaluated in place, CFL timestep and boundary inflow as
//                   per-block partials. In the second rk2 substep the kernel
//                   also applies friction, the update, fix_negative_cells and
//                   the final saxpy, since dt is already known.
//   de_update       friction + update + fix_negative_cells
//                   (sw_domain_orig.c:2805-2838, quantity.c:772-822,
//                   sw_domain_orig.c:2890-2910).
//   de_reduce       fixed-order reduction of the per-block partials.
// No atomics touch the state, so every run is bitwise reproducible.
#include "de_device.h"
#include "de_pow.h"

// centroid state; no __restrict__: the Euler update writes it in place
struct DeState {
    const double *w;
    const double *uh;
    const double *vh;
};

// ---------------------------------------------------------------------------
// extrapolation
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(DE_BLOCK)
de_extrapolate(DeMesh m, DeParams p, DeState s, const double *__restrict__ z, DeEdges e,
               unsigned char *__restrict__ flags, double *__restrict__ part_mass,
               DeControl *ctl)
{
    __shared__ double red[DE_BLOCK / 64];
    if (ctl->done) return;
    int *err = &ctl->status;
    const int N = m.N;
    const int k = xcd_block() * DE_BLOCK + threadIdx.x;
    double mass = 0.0;

    if (k < N) {
        const double mah = p.mah;
        // Loads are issued in three unconditional rounds (own cell, the three
        // surrogate neighbours with their connectivity, the neighbours'
        // neighbours' heights) so that each round's gathers overlap.
        // round 1: own cell
        int v[3], sur[3];
#pragma unroll
        for (int i = 0; i < 3; i++) v[i] = m.nbr[i * N + k];
        const double wk = s.w[k], zk = z[k], uhk = s.uh[k], vhk = s.vh[k];
        int nbound = 0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            sur[i] = v[i] >= 0 ? (v[i] >> 2) : k;
            nbound += v[i] < 0;
        }
        double dxv[3], dyv[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            dxv[i] = m.dxv[i * N + k];
            dyv[i] = m.dyv[i * N + k];
        }
        // round 2: surrogate neighbours
        double ws[3], zs[3], uhs[3], vhs[3], xs[3], ys[3];
        int vs[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int j = sur[i];
            ws[i] = s.w[j];
            zs[i] = z[j];
            uhs[i] = s.uh[j];
            vhs[i] = s.vh[j];
            xs[i] = m.cx[j];
            ys[i] = m.cy[j];
#pragma unroll
            for (int q = 0; q < 3; q++) vs[i][q] = m.nbr[q * N + j];
        }
        // round 3: heights of the neighbours' neighbours
        double h2[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const int t = vs[i][q] >= 0 ? (vs[i][q] >> 2) : sur[i];
                h2[i][q] = de_height(s.w[t], z[t]);
            }

        // protect + velocity of triangle k (sw_domain_orig.c:1245-1276, :2014-2064)
        const bool clamp = de_clamped(wk, zk, mah);
        DeCell c;
        c.w = clamp ? zk : wk;
        c.h = de_height(c.w, zk);
        double hs[3];
#pragma unroll
        for (int i = 0; i < 3; i++) hs[i] = de_height(ws[i], zs[i]);
        bool dry_k = true;
#pragma unroll
        for (int i = 0; i < 3; i++) dry_k = dry_k && (v[i] < 0 || hs[i] < mah);
        const bool keep = (c.h > mah) && !dry_k;
        c.u = 0.0;
        c.v = 0.0;
        if (keep) {
            const double inv = 1.0 / c.h;
            c.u = uhk * inv;
            c.v = vhk * inv;
        }
        if (clamp) mass = (zk - wk) * m.area[k];
        flags[k] = (clamp ? DE_F_CLAMP : 0u) | (keep ? DE_F_KEEP : 0u);

        // the same for the surrogate neighbours
        DeCell nb[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const bool cl = de_clamped(ws[i], zs[i], mah);
            nb[i].w = cl ? zs[i] : ws[i];
            nb[i].h = hs[i];
            bool dry = true;
#pragma unroll
            for (int q = 0; q < 3; q++) dry = dry && (vs[i][q] < 0 || h2[i][q] < mah);
            nb[i].u = 0.0;
            nb[i].v = 0.0;
            if (hs[i] > mah && !dry) {
                const double inv = 1.0 / hs[i];
                nb[i].u = uhs[i] * inv;
                nb[i].v = vhs[i] * inv;
            }
        }

        double ow[3], oh[3], ou[3], ov[3];
        bool constant = (nbound == 3);
        if (nbound <= 1) {
            const double dx1 = xs[1] - xs[0], dx2 = xs[2] - xs[0];
            const double dy1 = ys[1] - ys[0], dy2 = ys[2] - ys[0];
            const double area2 = dy2 * dx1 - dy1 * dx2;
            if (area2 <= 0.0) {
                constant = true;
            } else {
                const double a_tmp = 0.3, b_tmp = 0.1;
                const double c_tmp = 1.0 / (a_tmp - b_tmp);
                const double d_tmp = 1.0 - (c_tmp * a_tmp);
                const double hc = c.h, h0 = nb[0].h, h1 = nb[1].h, h2_ = nb[2].h;
                const double hmin = fmin(fmin(h0, fmin(h1, h2_)), hc);
                const double hmax = fmax(fmax(h0, fmax(h1, h2_)), hc);
                double hfactor = fmax(0.0, fmin(c_tmp * fmax(hmin, 0.0) / fmax(hc, 1.0e-06) + d_tmp,
                                                fmin(c_tmp * fmax(hc, 0.0) / fmax(hmax, 1.0e-06) + d_tmp,
                                                     1.0)));
                hfactor = fmin(1.2 * fmax(hmin - mah, 0.0) / (fmax(hmin, 0.0) + 1.0 * mah), hfactor);
                const double inv_area2 = 1.0 / area2;

                double beta = p.bwd + (p.bw - p.bwd) * hfactor;
                if (beta > 0.0) {
                    de_aux_edges(c.w, nb[0].w, nb[1].w, nb[2].w, dx1, dx2, dy1, dy2, inv_area2,
                                 dxv, dyv, beta, ow);
                    de_aux_edges(c.h, nb[0].h, nb[1].h, nb[2].h, dx1, dx2, dy1, dy2, inv_area2,
                                 dxv, dyv, beta, oh);
                } else {
                    ow[0] = ow[1] = ow[2] = c.w;
                    oh[0] = oh[1] = oh[2] = c.h;
                }
                beta = p.buhd + (p.buh - p.buhd) * hfactor;
                if (beta > 0.0) {
                    de_aux_edges(c.u, nb[0].u, nb[1].u, nb[2].u, dx1, dx2, dy1, dy2, inv_area2,
                                 dxv, dyv, beta, ou);
                } else {
                    ou[0] = ou[1] = ou[2] = c.u;
                }
                beta = p.bvhd + (p.bvh - p.bvhd) * hfactor;
                if (beta > 0.0) {
                    de_aux_edges(c.v, nb[0].v, nb[1].v, nb[2].v, dx1, dx2, dy1, dy2, inv_area2,
                                 dxv, dyv, beta, ov);
                } else {
                    ov[0] = ov[1] = ov[2] = c.v;
                }
            }
        } else if (nbound == 2) {
            // the first edge whose surrogate is not k itself (sw_domain_orig.c:2421-2439)
            const int i1 = (v[0] >= 0) ? 0 : ((v[1] >= 0) ? 1 : 2);
            if (v[i1] < 0) {
                atomicOr(err, DE_ERR_NO_NEIGHBOUR);
                constant = true;
            } else {
                const double dx1 = xs[i1] - m.cx[k];
                const double dy1 = ys[i1] - m.cy[k];
                const double area2 = dx1 * dx1 + dy1 * dy1;
                double dx2 = 1.0 / area2;
                const double dy2 = dx2 * dy1;
                dx2 *= dx1;
                de_line_edges(c.w, nb[i1].w, dx2, dy2, dxv, dyv, p.bw, ow);
                de_line_edges(c.h, nb[i1].h, dx2, dy2, dxv, dyv, p.bw, oh);
                de_line_edges(c.u, nb[i1].u, dx2, dy2, dxv, dyv, p.bw, ou);
                de_line_edges(c.v, nb[i1].v, dx2, dy2, dxv, dyv, p.bw, ov);
            }
        }
        if (constant) {
#pragma unroll
            for (int i = 0; i < 3; i++) {
                ow[i] = c.w;
                oh[i] = c.h;
                ou[i] = c.u;
                ov[i] = c.v;
            }
        }
#pragma unroll
        for (int i = 0; i < 3; i++) {
            e.we[i * N + k] = ow[i];
            e.he[i * N + k] = oh[i];
            e.uhe[i * N + k] = ou[i] * oh[i];
            e.vhe[i * N + k] = ov[i] * oh[i];
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
        part_mass[xcd_block()] = t;
    }
}

// Write the protect / momentum-zeroing outcome back into the centroid state
// (what the reference's distribute_to_vertices_and_edges leaves behind).
__global__ void __launch_bounds__(DE_BLOCK)
de_materialize(int N, double *__restrict__ w, double *__restrict__ uh, double *__restrict__ vh,
               const double *__restrict__ z, const unsigned char *__restrict__ flags)
{
    const int k = xcd_block() * DE_BLOCK + threadIdx.x;
    if (k >= N) return;
    const unsigned f = flags[k];
    if (f & DE_F_CLAMP) w[k] = z[k];
    if (!(f & DE_F_KEEP)) {
        uh[k] = 0.0;
        vh[k] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// friction + update + fix_negative_cells for one triangle
// ---------------------------------------------------------------------------
struct DeUpdated {
    double w, uh, vh;
};

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

__device__ __forceinline__ DeUpdated de_advance(double w, double uh, double vh, double z,
                                                double eta, double exw, double exuh,
                                                double exvh, double dt, const DeParams &p,
                                                int *err, unsigned long long *nneg)
{
    // _manning_friction_flat with eps = minimum_allowed_height (friction.py:64-72)
    double si_uh = 0.0, si_vh = 0.0;
    if (eta > p.mah) {
        const double h = w - z;
        if (h >= p.mah) {
            double S = -p.g * eta * eta * sqrt((uh * uh + vh * vh));
            S /= de_pow_7_3(h);  // correctly rounded pow(h, 7.0/3.0)
            si_uh += S * uh;
            si_vh += S * vh;
        }
    }
    DeUpdated r;
    r.w = de_update_one(w, exw, 0.0, dt, err);
    r.uh = de_update_one(uh, exuh, si_uh, dt, err);
    r.vh = de_update_one(vh, exvh, si_vh, dt, err);
    // _orig_fix_negative_cells (every triangle is a full triangle here)
    if (r.w - z < 0.0) {
        atomicAdd(nneg, 1ull);
        r.w = z;
        r.uh = 0.0;
        r.vh = 0.0;
    }
    return r;
}

// ---------------------------------------------------------------------------
// flux
// ---------------------------------------------------------------------------
// SUBSTEP 0: writes explicit updates; SUBSTEP 1: advances with dt and
// combines with the backup state (rk2 final stage).
template <int SUBSTEP>
__global__ void __launch_bounds__(DE_BLOCK)
de_flux(DeMesh m, DeParams p, DeState s, const double *__restrict__ z,
        const double *__restrict__ eta, DeEdges ec, DeBoundary bc,
        const unsigned char *__restrict__ flags, double *__restrict__ exw,
        double *__restrict__ exuh, double *__restrict__ exvh, double *__restrict__ bw,
        double *__restrict__ buh, double *__restrict__ bvh, DeControl *ctl,
        double *__restrict__ part_dt, double *__restrict__ part_bf,
        unsigned long long *__restrict__ nneg)
{
    __shared__ double red_dt[DE_BLOCK / 64];
    __shared__ double red_bf[DE_BLOCK / 64];
    if (ctl->done) return;
    int *err = &ctl->status;
    const double dt = SUBSTEP ? ctl->dt : 0.0;
    const int N = m.N;
    const int k = xcd_block() * DE_BLOCK + threadIdx.x;
    double dtmin = 1.0e100, bflux = 0.0;

    if (k < N) {
        const double g = p.g, eps = p.eps;
        const unsigned f = flags[k];
        const double zc = z[k];
        const double wc = (f & DE_F_CLAMP) ? zc : s.w[k];
        const double uhc = (f & DE_F_KEEP) ? s.uh[k] : 0.0;
        const double vhc = (f & DE_F_KEEP) ? s.vh[k] : 0.0;
        const double hc = de_height(wc, zc);
        double sw = 0.0, sx = 0.0, sy = 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int ki = i * N + k;
            const double ql0 = ec.we[ki], hle = ec.he[ki];
            const double ql1 = ec.uhe[ki], ql2 = ec.vhe[ki];
            const double zl = ql0 - hle;
            const double n1 = m.nx[ki], n2 = m.ny[ki];
            const int v = m.nbr[ki];
            double qr0, qr1, qr2, zr, hre;
            if (v >= 0) {
                const int nm = (v & 3) * N + (v >> 2);
                qr0 = ec.we[nm];
                qr1 = ec.uhe[nm];
                qr2 = ec.vhe[nm];
                hre = ec.he[nm];
                zr = qr0 - hre;
            } else {
                const int j = -v - 1;
                const int kind = bc.kind[j];
                if (kind == 0) {          // reflective
                    const double r1 = -ql1 * n1 - ql2 * n2;
                    const double r2 = -ql1 * n2 + ql2 * n1;
                    qr0 = ql0;
                    qr1 = n1 * r1 - n2 * r2;
                    qr2 = n2 * r1 + n1 * r2;
                } else if (kind == 1) {   // transmissive
                    qr0 = ql0;
                    qr1 = ql1;
                    qr2 = ql2;
                } else if (kind == 2) {   // Dirichlet
                    qr0 = bc.val[j];
                    qr1 = bc.val[bc.M + j];
                    qr2 = bc.val[2 * bc.M + j];
                } else {                  // transmissive from centroid values
                    qr0 = wc;
                    qr1 = uhc;
                    qr2 = vhc;
                }
                zr = zl;
                hre = fmax(qr0 - zr, 0.0);
            }
            const double z_half = fmax(zl, zr);
            const double h_left = fmax(hle + zl - z_half, 0.0);
            const double h_right = fmax(hre + zr - z_half, 0.0);
            double ef[3], speed, pflux;
            de_flux_central(ql0, ql1, ql2, qr0, qr1, qr2, h_left, h_right, hle, hre, n1, n2, eps,
                            z_half, g, ef, speed, pflux);
            const double length = m.len[ki];
            ef[0] *= length;
            ef[1] *= length;
            ef[2] *= length;
            const double pg =
                length * (-g * 0.5 * (h_left * h_left - hle * hle - (hle + hc) * (zl - zc)) + pflux);
            sw += -ef[0];
            sx += -ef[1];
            sy += -ef[2];
            if (v < 0) bflux += -ef[0];
            sx -= n1 * pg;
            sy -= n2 * pg;
            if (SUBSTEP == 0 && speed > eps) {
                const double t = 1.0 / fmax(speed, eps);
                dtmin = fmin(dtmin, m.radius[k] * t);
            }
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
            const DeUpdated r = de_advance(wc, uhc, vhc, zc, eta[k], sw, sx, sy, dt, p, err, nneg);
            // saxpy_centroid_values(0.5, 0.5) with the backup (quantity.c:752-769)
            bw[k] = 0.5 * r.w + 0.5 * bw[k];
            buh[k] = 0.5 * r.uh + 0.5 * buh[k];
            bvh[k] = 0.5 * r.vh + 0.5 * bvh[k];
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
        const int lb = xcd_block();
        part_dt[lb] = a;
        part_bf[lb] = b;
    }
}

// friction + update + fix_negative_cells from the explicit updates of substep 0
__global__ void __launch_bounds__(DE_BLOCK)
de_update(int N, DeParams p, DeState s, const double *__restrict__ z,
          const double *__restrict__ eta, const unsigned char *__restrict__ flags,
          const double *__restrict__ exw, const double *__restrict__ exuh,
          const double *__restrict__ exvh, DeControl *ctl, double *ow, double *ouh, double *ovh,
          unsigned long long *__restrict__ nneg)
{
    if (ctl->done) return;
    const int k = xcd_block() * DE_BLOCK + threadIdx.x;
    if (k >= N) return;
    const double dt = ctl->dt;
    int *err = &ctl->status;
    const unsigned f = flags[k];
    const double zc = z[k];
    const double wc = (f & DE_F_CLAMP) ? zc : s.w[k];
    const double uhc = (f & DE_F_KEEP) ? s.uh[k] : 0.0;
    const double vhc = (f & DE_F_KEEP) ? s.vh[k] : 0.0;
    const DeUpdated r = de_advance(wc, uhc, vhc, zc, eta[k], exw[k], exuh[k], exvh[k], dt, p, err, nneg);
    ow[k] = r.w;
    ouh[k] = r.uh;
    ovh[k] = r.vh;
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
// Fixed-order reduction of the per-block partials into the control block,
// one block of 1024 threads. mode 0: flux_dt, bflux0, mass (first substep);
// 1: bflux1 (second substep); 2: mass (distribute); 3: volume.
__global__ void __launch_bounds__(1024)
de_reduce(int nparts, const double *__restrict__ part_dt, const double *__restrict__ part_bf,
          const double *__restrict__ part_mass, DeControl *ctl, int mode)
{
    __shared__ double s0[16], s1[16], s2[16];
    if (mode < 2 && ctl->done) return;
    double a = 1.0e100, b = 0.0, c = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 1024) {
        if (part_dt) a = fmin(a, part_dt[i]);
        if (part_bf) b += part_bf[i];
        if (part_mass) c += part_mass[i];
    }
    a = de_wave_min(a);
    b = de_wave_sum(b);
    c = de_wave_sum(c);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        s0[wv] = a;
        s1[wv] = b;
        s2[wv] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < 16; i++) {
            s0[0] = fmin(s0[0], s0[i]);
            s1[0] += s1[i];
            s2[0] += s2[i];
        }
        if (mode == 0) {
            ctl->flux_dt = s0[0];
            ctl->bflux0 = s1[0];
            ctl->mass = s2[0];
        } else if (mode == 1) {
            ctl->bflux1 = s1[0];
        } else if (mode == 2) {
            ctl->mass = s2[0];
        } else {
            ctl->volume = s2[0];
        }
    }
}

// Generic_Domain.update_timestep (generic_domain.py:2349-2415) on the device:
// CFL and max-timestep limit, recorded extremes, the small-step counter with
// its first-order switch and its exception (reported as DE_ERR_SMALL_TIMESTEP),
// and the clamps to the final time and the next yield time.
__global__ void de_timestep(DeControl *c)
{
    if (threadIdx.x != 0 || blockIdx.x != 0 || c->done) return;
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
// relative_time += dt, timestep history, and the yield / final-time test of
// _evolve_base (generic_domain.py:1855-1895) that turns the next steps into no-ops.
__global__ void de_finish_step(DeControl *c, double *hist)
{
    if (threadIdx.x != 0 || blockIdx.x != 0 || c->done) return;
    const double dt = c->dt;
    if (c->rk2) c->integral = c->integral + 0.5 * dt * (c->bflux0 + c->bflux1);
    else c->integral = c->integral + dt * c->bflux0;
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

// host-controlled steps: the host computed dt
__global__ void de_set_dt(DeControl *c, double dt)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) c->dt = dt;
}

// host-controlled steps: boundary flux integral after the update
__global__ void de_integrate(DeControl *c, int rk2)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (rk2) c->integral = c->integral + 0.5 * c->dt * (c->bflux0 + c->bflux1);
    else c->integral = c->integral + c->dt * c->bflux0;
}

// time-control parameters and state for a run of device-controlled steps
struct DeBatchSetup {
    double t, rec_min, rec_max, cfl, max_dt, min_dt, finaltime, yieldtime, fixed_flux_dt, eps;
    int smallsteps, order, default_order, max_smallsteps, hist_cap, rk2;
};

__global__ void de_setup_batch(DeControl *c, DeBatchSetup b)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
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
    c->rk2 = b.rk2;
    c->nsteps = 0;
    c->done = 0;
}

// leave device-controlled mode: steps are live again
__global__ void de_clear_done(DeControl *c)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) c->done = 0;
}

// per-block partial of sum(area * max(w - z, 0)) over the current state
__global__ void __launch_bounds__(DE_BLOCK)
de_volume(int N, const double *__restrict__ w, const double *__restrict__ z,
          const double *__restrict__ area, double *__restrict__ part)
{
    __shared__ double red[DE_BLOCK / 64];
    const int k = xcd_block() * DE_BLOCK + threadIdx.x;
    double v = 0.0;
    if (k < N) v = area[k] * de_height(w[k], z[k]);
    v = de_wave_sum(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) red[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = red[0];
        for (int i = 1; i < DE_BLOCK / 64; i++) t += red[i];
        part[xcd_block()] = t;
    }
}

template __global__ void de_flux<0>(DeMesh, DeParams, DeState, const double *, const double *,
                                    DeEdges, DeBoundary, const unsigned char *, double *,
                                    double *, double *, double *, double *, double *,
                                    DeControl *, double *, double *, unsigned long long *);
template __global__ void de_flux<1>(DeMesh, DeParams, DeState, const double *, const double *,
                                    DeEdges, DeBoundary, const unsigned char *, double *,
                                    double *, double *, double *, double *, double *,
                                    DeControl *, double *, double *, unsigned long long *);

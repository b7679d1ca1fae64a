// CPU driver of the generic DE cell math (anuga_core_amd/csrc/de_adjoint.h),
// for tests only: a fixed-dt DE1 forward over the library's layouts in the
// caller's triangle order, and its discrete adjoint, so that both can be
// checked on the CPU against the oracle (forward, bit for bit) and against
// finite differences (adjoint).
#include <string.h>

#include <vector>

#include "de_adjoint.h"

namespace {

double g_freeze = 1.0e-10;   // AdjParams::freeze of the runs (the library's value)

struct Run {
    AdjView m;
    int N;
    // one DE1 step from S0: stage states, flags, edge values, explicit updates
    std::vector<double> S1, E0, E1, ex0, ex1;
    std::vector<unsigned char> f0, f1;
    void resize()
    {
        S1.resize(4 * N);
        E0.resize(12 * N);
        E1.resize(12 * N);
        ex0.resize(3 * N);
        ex1.resize(3 * N);
        f0.resize(N);
        f1.resize(N);
    }
    void stage(const double *S, std::vector<unsigned char> &f, std::vector<double> &E,
               std::vector<double> &ex, double *out)
    {
        for (int k = 0; k < N; k++) f[k] = (unsigned char)adj_flags(m, S, k);
        for (int k = 0; k < N; k++) adj_fwd_extrap(m, S, f.data(), k, E.data());
        for (int k = 0; k < N; k++) {
            double e[3];
            adj_fwd_flux(m, S, f.data(), E.data(), k, e);
            ex[k] = e[0];
            ex[N + k] = e[1];
            ex[2 * N + k] = e[2];
        }
        for (int k = 0; k < N; k++) {
            const double e[3] = {ex[k], ex[N + k], ex[2 * N + k]};
            double r[3];
            adj_fwd_advance(m, S, f.data(), e, k, r);
            out[4 * k] = r[0];
            out[4 * k + 1] = r[1];
            out[4 * k + 2] = r[2];
            out[4 * k + 3] = S[4 * k + 3];
        }
    }
    // S (in) -> Snew (out)
    void step(const double *S, double *Snew)
    {
        std::vector<double> S2(4 * N);
        stage(S, f0, E0, ex0, S1.data());
        stage(S1.data(), f1, E1, ex1, S2.data());
        for (int k = 0; k < N; k++) {
            for (int c = 0; c < 3; c++) Snew[4 * k + c] = 0.5 * S2[4 * k + c] + 0.5 * S[4 * k + c];
            Snew[4 * k + 3] = S[4 * k + 3];
        }
    }
    // adjoint of one stage: lam_out (of its result) -> lamS (of its input), zbar, nbar
    void stage_adjoint(const double *S, const std::vector<unsigned char> &f,
                       const std::vector<double> &E, const std::vector<double> &ex,
                       const double *lam_out, double *lamS, double *zbar, double *nbar)
    {
        std::vector<double> lam_ex(3 * N), lamE(12 * N, 0.0);
        for (int k = 0; k < N; k++) {
            const double lo[3] = {lam_out[k], lam_out[N + k], lam_out[2 * N + k]};
            adj_bwd_advance(m, S, f.data(), ex.data(), ex.data() + N, ex.data() + 2 * N, lo, k,
                            lamS, zbar, nbar, lam_ex.data());
        }
        for (int k = 0; k < N; k++)
            adj_bwd_flux_rev(m, S, f.data(), E.data(), lam_ex.data(), k, lamE.data(), lamS, zbar);
        std::vector<double> X(16 * (size_t)N);
        for (int k = 0; k < N; k++) adj_bwd_extrap_rev(m, S, f.data(), lamE.data(), k, X.data());
        for (int k = 0; k < N; k++) adj_gather_extrap(m, X.data(), k, lamS, zbar);
    }
    // lam (of the step's result, [3][N]) -> lam (of its start), given the start S
    void step_adjoint(const double *S, double *lam, double *zbar, double *nbar)
    {
        std::vector<double> tmp(4 * N);
        step(S, tmp.data());   // recompute the stages of this step
        std::vector<double> l2(3 * N), l1(3 * N, 0.0), l0(3 * N);
        for (int i = 0; i < 3 * N; i++) {
            l2[i] = 0.5 * lam[i];
            l0[i] = 0.5 * lam[i];
        }
        stage_adjoint(S1.data(), f1, E1, ex1, l2.data(), l1.data(), zbar, nbar);
        stage_adjoint(S, f0, E0, ex0, l1.data(), l0.data(), zbar, nbar);
        memcpy(lam, l0.data(), 3 * N * sizeof(double));
    }
};

Run make_run(int N, int M, const int *nb, const double *cen, const double *gx, const double *nx,
             const double *ny, const double *len, const double *area, const double *eta,
             const signed char *bkind, const double *bval, const double *prm, double dt)
{
    Run r;
    r.N = N;
    r.m.N = N;
    r.m.M = M;
    r.m.nb = nb;
    r.m.cen = cen;
    r.m.gx = gx;
    r.m.gx_stride = 6;
    r.m.nx = nx;
    r.m.ny = ny;
    r.m.len = len;
    r.m.area = area;
    r.m.eta = eta;
    r.m.bkind = bkind;
    r.m.bval = bval;
    r.m.p = AdjParams{prm[0], prm[1], prm[2], prm[3], prm[4], prm[5], prm[6], prm[7], prm[8],
                      g_freeze};
    r.m.dt = dt;
    r.resize();
    return r;
}

}  // namespace

extern "C" {

// limiter jumps below `freeze` hold the limiter's factor in the adjoint
void harness_set_freeze(double freeze) { g_freeze = freeze; }

// nsteps fixed-dt DE1 steps of S ([N][4], in place); stage of the gauge
// triangles after every obs_every-th step into rec ([nsteps / obs_every][G]).
void harness_forward(int N, int M, const int *nb, const double *cen, const double *gx,
                     const double *nx, const double *ny, const double *len, const double *area,
                     const double *eta, const signed char *bkind, const double *bval,
                     const double *prm, double dt, int nsteps, double *S, int G,
                     const int *gauges, int obs_every, double *rec)
{
    Run r = make_run(N, M, nb, cen, gx, nx, ny, len, area, eta, bkind, bval, prm, dt);
    std::vector<double> T(4 * N);
    for (int s = 1; s <= nsteps; s++) {
        r.step(S, T.data());
        memcpy(S, T.data(), 4 * N * sizeof(double));
        if (G > 0 && s % obs_every == 0)
            for (int g = 0; g < G; g++) rec[(s / obs_every - 1) * G + g] = S[4 * gauges[g]];
    }
}

// Gradient of L = sum(dL/dstage * stage at the gauges) (dL given per record)
// with respect to the friction (nbar), the bed (zbar) and the initial state
// (lam, [3][N]), by the discrete adjoint of harness_forward from S0.
void harness_adjoint(int N, int M, const int *nb, const double *cen, const double *gx,
                     const double *nx, const double *ny, const double *len, const double *area,
                     const double *eta, const signed char *bkind, const double *bval,
                     const double *prm, double dt, int nsteps, const double *S0, int G,
                     const int *gauges, int obs_every, const double *dL, double *nbar,
                     double *zbar, double *lam)
{
    Run r = make_run(N, M, nb, cen, gx, nx, ny, len, area, eta, bkind, bval, prm, dt);
    std::vector<std::vector<double>> states(nsteps + 1, std::vector<double>(4 * N));
    memcpy(states[0].data(), S0, 4 * N * sizeof(double));
    for (int s = 1; s <= nsteps; s++) r.step(states[s - 1].data(), states[s].data());
    memset(lam, 0, 3 * N * sizeof(double));
    memset(nbar, 0, N * sizeof(double));
    memset(zbar, 0, N * sizeof(double));
    for (int s = nsteps; s >= 1; s--) {
        if (G > 0 && s % obs_every == 0)
            for (int g = 0; g < G; g++) lam[gauges[g]] += dL[(s / obs_every - 1) * G + g];
        r.step_adjoint(states[s - 1].data(), lam, zbar, nbar);
    }
}

// the extrapolation's vector-Jacobian product per stencil slot (X, [16][N])
// for the edge-value adjoints wt ([4][3][N]), by forward-mode duals and by
// the reverse sweep
void harness_extrap_products(int N, int M, const int *nb, const double *cen, const double *gx,
                             const double *nx, const double *ny, const double *len,
                             const double *area, const double *eta, const signed char *bkind,
                             const double *bval, const double *prm, const double *S,
                             const double *wt, double *X_dual, double *X_rev)
{
    Run r = make_run(N, M, nb, cen, gx, nx, ny, len, area, eta, bkind, bval, prm, 0.1);
    std::vector<unsigned char> f(N);
    for (int k = 0; k < N; k++) f[k] = (unsigned char)adj_flags(r.m, S, k);
    for (int k = 0; k < N; k++) {
        adj_bwd_extrap(r.m, S, f.data(), wt, k, X_dual);
        adj_bwd_extrap_rev(r.m, S, f.data(), wt, k, X_rev);
    }
}


// the flux's vector-Jacobian product both ways on the edge values of S:
// wt [3][N] weights the explicit updates; out_* receive lamE [4][3][N] then
// lamS [3][N] then zbar [N] (16 N values each)
void harness_flux_products(int N, int M, const int *nb, const double *cen, const double *gx,
                           const double *nx, const double *ny, const double *len,
                           const double *area, const double *eta, const signed char *bkind,
                           const double *bval, const double *prm, const double *S,
                           const double *wt, double *out_dual, double *out_rev)
{
    Run r = make_run(N, M, nb, cen, gx, nx, ny, len, area, eta, bkind, bval, prm, 0.1);
    std::vector<unsigned char> f(N);
    for (int k = 0; k < N; k++) f[k] = (unsigned char)adj_flags(r.m, S, k);
    std::vector<double> E(12 * (size_t)N);
    for (int k = 0; k < N; k++) adj_fwd_extrap(r.m, S, f.data(), k, E.data());
    const size_t n = (size_t)N;
    for (int k = 0; k < N; k++) {
        adj_bwd_flux(r.m, S, f.data(), E.data(), wt, k, out_dual, out_dual + 12 * n,
                     out_dual + 15 * n);
        adj_bwd_flux_rev(r.m, S, f.data(), E.data(), wt, k, out_rev, out_rev + 12 * n,
                         out_rev + 15 * n);
    }
}
}  // extern "C"

extern "C" {

// pieces of one stage, for isolating the vector-Jacobian products:
// part 0: S -> E (extrapolation; flags from S); part 1: (S, E) -> ex (flux,
// E held fixed, S through the centroid terms); part 2: E -> ex (flux through
// the edge values). out receives sum(wt * output); grads: lamS [3][N], zbar,
// lamE [4][3][N].
double harness_part(int N, int M, const int *nb, const double *cen, const double *gx,
                    const double *nx, const double *ny, const double *len, const double *area,
                    const double *eta, const signed char *bkind, const double *bval,
                    const double *prm, int part, const double *S, const double *E,
                    const double *wt, double *lamS, double *zbar, double *lamE)
{
    Run r = make_run(N, M, nb, cen, gx, nx, ny, len, area, eta, bkind, bval, prm, 0.1);
    std::vector<unsigned char> f(N);
    for (int k = 0; k < N; k++) f[k] = (unsigned char)adj_flags(r.m, S, k);
    double val = 0.0;
    if (part == 0) {
        std::vector<double> Eo(12 * N);
        for (int k = 0; k < N; k++) adj_fwd_extrap(r.m, S, f.data(), k, Eo.data());
        for (int i = 0; i < 12 * N; i++) val += wt[i] * Eo[i];
        if (lamS) {
            std::vector<double> X(16 * (size_t)N);
            for (int k = 0; k < N; k++) adj_bwd_extrap(r.m, S, f.data(), wt, k, X.data());
            for (int k = 0; k < N; k++) adj_gather_extrap(r.m, X.data(), k, lamS, zbar);
        }
    } else {
        for (int k = 0; k < N; k++) {
            double e[3];
            adj_fwd_flux(r.m, S, f.data(), E, k, e);
            val += wt[k] * e[0] + wt[N + k] * e[1] + wt[2 * N + k] * e[2];
        }
        if (lamS)
            for (int k = 0; k < N; k++)
                adj_bwd_flux_rev(r.m, S, f.data(), E, wt, k, lamE, lamS, zbar);
    }
    return val;
}

}  // extern "C"

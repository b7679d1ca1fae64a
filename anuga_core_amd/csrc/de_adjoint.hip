// Discrete adjoint of the FP64 DE1 step with a fixed timestep, on the device.
//
// One thread per triangle, as the forward. Each kernel applies one
// vector-Jacobian product of de_adjoint.h to its triangle. The adjoint of a
// gather is a scatter; it is kept deterministic: the update and flux kernels
// add only into their own triangle's entries, except the flux's adjoint of a
// neighbour's edge value, an entry that receives exactly two contributions
// into a zeroed array (a + b in either order); the extrapolation stores its
// contributions per stencil slot and a second kernel sums each triangle's in
// a fixed order. Gradients are therefore the same bits on every run. The
// forward quantities each product needs (flags, edge values, explicit
// updates, the stage state) come from the production kernels, recomputed
// step by step from checkpoints.
#include "de_adjoint.h"

__device__ __forceinline__ int adj_tid() { return blockIdx.x * blockDim.x + threadIdx.x; }

// adjoint of the update (friction, update, fix_negative) of one stage;
// lam_out is scaled by `scale` (0.5: the final saxpy of the step)
__global__ void __launch_bounds__(DE_BLOCK)
adj_k_advance(AdjView m, const double *__restrict__ S, const unsigned char *__restrict__ fl,
              const double *__restrict__ exw, const double *__restrict__ exuh,
              const double *__restrict__ exvh, const double *__restrict__ lam_out, double scale,
              double *lamS, double *zbar, double *nbar, double *__restrict__ lam_ex)
{
    const int k = adj_tid();
    if (k >= m.N) return;
    const double lo[3] = {scale * lam_out[k], scale * lam_out[m.N + k],
                          scale * lam_out[2 * (size_t)m.N + k]};
    adj_bwd_advance(m, S, fl, exw, exuh, exvh, lo, k, lamS, zbar, nbar, lam_ex);
}

__global__ void __launch_bounds__(DE_BLOCK)
adj_k_flux(AdjView m, const double *__restrict__ S, const unsigned char *__restrict__ fl,
           const double *__restrict__ E, const double *__restrict__ lam_ex, double *lamE,
           double *lamS, double *zbar)
{
    const int k = adj_tid();
    if (k >= m.N) return;
    adj_bwd_flux_rev(m, S, fl, E, lam_ex, k, lamE, lamS, zbar);
}

__global__ void __launch_bounds__(DE_BLOCK)
adj_k_extrap(AdjView m, const double *__restrict__ S, const unsigned char *__restrict__ fl,
             const double *__restrict__ lamE, double *__restrict__ X)
{
    const int k = adj_tid();
    if (k >= m.N) return;
    adj_bwd_extrap_rev(m, S, fl, lamE, k, X);
}

// the extrapolation adjoints stored per stencil slot, summed per triangle in
// a fixed order (no two threads add to one entry)
__global__ void __launch_bounds__(DE_BLOCK)
adj_k_extrap_gather(AdjView m, const double *__restrict__ X, double *lamS, double *zbar)
{
    const int t = adj_tid();
    if (t >= m.N) return;
    adj_gather_extrap(m, X, t, lamS, zbar);
}

// out = a * in over n entries
__global__ void __launch_bounds__(DE_BLOCK)
adj_k_scale(int n, double a, const double *__restrict__ in, double *__restrict__ out)
{
    const int i = adj_tid();
    if (i < n) out[i] = a * in[i];
}

// stage of the gauge triangles into a record row
__global__ void adj_k_gauges(int G, const int *__restrict__ gauges, const DeQ *__restrict__ S,
                             double *__restrict__ row)
{
    const int g = adj_tid();
    if (g < G) row[g] = S[gauges[g]].w;
}

// add the loss's sensitivity to the gauges' stage into the stage adjoint
// (gauges are distinct)
__global__ void adj_k_gauge_seed(int G, const int *__restrict__ gauges,
                                 const double *__restrict__ row, double *lam)
{
    const int g = adj_tid();
    if (g < G) lam[gauges[g]] += row[g];
}

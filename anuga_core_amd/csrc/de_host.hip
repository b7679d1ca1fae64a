// Host runtime and C ABI of the DE step (see include/anuga_hip.h).
//
// The handle owns all device buffers. Triangles are renumbered once at
// creation (Reverse-Cuthill-McKee on the triangle adjacency) so that neighbour
// gathers stay within a narrow index window; the permutation is applied on
// upload and undone on download, so callers only ever see the reference's
// order. Per-edge and per-triangle arrays are stored structure-of-arrays
// ([edge][triangle]) for coalesced 64-lane loads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "anuga_hip.h"
#include "de_kernels.hip"

namespace {

thread_local std::string g_create_error;

struct KernelTimer {
    std::vector<hipEvent_t> start, stop;
    int pending = 0;
    double total_ms = 0.0;
    int64_t launches = 0;
};

const char *kTimerNames[] = {"extrapolate", "flux0", "update", "flux1"};
enum { T_EXTRAP = 0, T_FLUX0, T_UPDATE, T_FLUX1, T_COUNT };
const int kTimerRing = 64;

}  // namespace

struct anuga_domain {
    int N = 0, M = 0, nblocks = 0;
    anuga_params P{};
    hipStream_t stream = nullptr;
    std::vector<int> perm;   // perm[new] = old
    std::vector<int> iperm;  // iperm[old] = new

    // static mesh records, internal order
    int4 *nb = nullptr;
    double2 *cen = nullptr;
    void *gx = nullptr, *gf = nullptr;   // DeGeoX64/F64 or DeGeoX32/F32
    double *area = nullptr;
    double *nx = nullptr, *ny = nullptr, *len = nullptr, *radius = nullptr;  // FP64 flux geometry
    bool mixed = false;
    std::vector<DeGeoF32> gf32;   // host copy of the mixed flux geometry (bed offsets change)
    // quantities
    double *eta = nullptr;
    DeQ *A = nullptr;   // step state / rk2 backup: {w, uh, vh, z}
    DeQ *B = nullptr;   // rk2 substep state
    void *E = nullptr;  // 3N DeEdge4<double> or DeEdge4<float>
    unsigned char *flags = nullptr;
    double *ex[3] = {nullptr, nullptr, nullptr};
    signed char *bkind = nullptr;
    double *bval = nullptr;
    double *part = nullptr;   // [3][nblocks]: dt, bflux, mass
    DeControl *ctl = nullptr; // device step control
    unsigned long long *dneg = nullptr;
    DeControl *hctl = nullptr;          // pinned host mirror of *ctl
    unsigned long long *hneg = nullptr; // pinned host mirror of *dneg
    double *hist = nullptr;             // device timestep history of a step run
    int64_t hist_cap = 0;
    bool done_on_device = false;        // a step run may have left ctl->done set
    hipGraph_t graph = nullptr;         // one captured device-controlled step
    hipGraphExec_t graph_exec = nullptr;
    bool use_graph = true;

    int substate = 0;         // 0: at a step boundary (state in A); 1: rk2 substep state in B
    bool boundary_set = false;
    unsigned long long neg_seen = 0;
    std::string err;

    bool timing = false;
    KernelTimer timers[T_COUNT];
};

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
static int fail(anuga_domain *d, int code, const std::string &msg)
{
    if (d) d->err = msg;
    else g_create_error = msg;
    return code;
}

#define HIPCHK(d, call)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail((d), ANUGA_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
static hipError_t dalloc(T **p, size_t n)
{
    return hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
}

// Reverse Cuthill-McKee over the triangle adjacency (interior edges). Each
// connected component starts from a pseudo-peripheral triangle (George-Liu);
// neighbours are queued by increasing degree, ties by index.
static std::vector<int> rcm_order(int N, const int64_t *neighbours)
{
    std::vector<int> deg(N, 0);
    for (int k = 0; k < N; k++)
        for (int i = 0; i < 3; i++)
            if (neighbours[3 * (int64_t)k + i] >= 0) deg[k]++;
    auto nb = [&](int k, int i) { return (int)neighbours[3 * (int64_t)k + i]; };

    std::vector<int> level(N, -1);
    std::vector<char> placed(N, 0);
    std::vector<int> order;
    order.reserve(N);
    std::vector<int> queue;
    queue.reserve(N);

    auto bfs_last_level = [&](int root, int &ecc) {
        // BFS restricted to the component; returns the min-degree node of the last level
        queue.clear();
        queue.push_back(root);
        level[root] = 0;
        size_t head = 0;
        int last = root;
        ecc = 0;
        while (head < queue.size()) {
            int u = queue[head++];
            for (int i = 0; i < 3; i++) {
                int v = nb(u, i);
                if (v >= 0 && level[v] < 0) {
                    level[v] = level[u] + 1;
                    queue.push_back(v);
                }
            }
        }
        ecc = level[queue.back()];
        int best = queue.back();
        for (size_t q = queue.size(); q-- > 0;) {
            int u = queue[q];
            if (level[u] != ecc) break;
            if (deg[u] < deg[best] || (deg[u] == deg[best] && u < best)) best = u;
        }
        for (int u : queue) level[u] = -1;
        last = best;
        return last;
    };

    for (int seed = 0; seed < N; seed++) {
        if (placed[seed]) continue;
        int root = seed, ecc = 0;
        int cand = bfs_last_level(root, ecc);
        for (int it = 0; it < 8; it++) {
            int ecc2 = 0;
            int c2 = bfs_last_level(cand, ecc2);
            if (ecc2 <= ecc) break;
            root = cand;
            cand = c2;
            ecc = ecc2;
        }
        root = cand;
        // Cuthill-McKee from root
        size_t head = order.size();
        order.push_back(root);
        placed[root] = 1;
        while (head < order.size()) {
            int u = order[head++];
            int cs[3], nc = 0;
            for (int i = 0; i < 3; i++) {
                int v = nb(u, i);
                if (v >= 0 && !placed[v]) {
                    placed[v] = 1;
                    cs[nc++] = v;
                }
            }
            std::sort(cs, cs + nc, [&](int a, int b) {
                return deg[a] < deg[b] || (deg[a] == deg[b] && a < b);
            });
            for (int q = 0; q < nc; q++) order.push_back(cs[q]);
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

static DeParams dev_params(const anuga_params &p)
{
    DeParams q;
    q.g = p.g;
    q.eps = p.epsilon;
    q.mah = p.minimum_allowed_height;
    q.bw = p.beta_w;
    q.bwd = p.beta_w_dry;
    q.buh = p.beta_uh;
    q.buhd = p.beta_uh_dry;
    q.bvh = p.beta_vh;
    q.bvhd = p.beta_vh_dry;
    return q;
}

static int check_params(anuga_domain *d, const anuga_params *p)
{
    if (!p) return fail(d, ANUGA_E_ARG, "params is NULL");
    if (p->low_froude != 0)
        return fail(d, ANUGA_E_ARG, "low_froude != 0 is not supported on the device path");
    if (p->extrapolate_velocity_second_order != 1)
        return fail(d, ANUGA_E_ARG, "extrapolate_velocity_second_order must be 1 (DE algorithms)");
    if (p->timestep_fluxcalls != 1 && p->timestep_fluxcalls != 2)
        return fail(d, ANUGA_E_ARG, "timestep_fluxcalls must be 1 (euler) or 2 (rk2)");
    if (p->precision != ANUGA_PRECISION_FP64 && p->precision != ANUGA_PRECISION_MIXED)
        return fail(d, ANUGA_E_ARG, "unknown precision");
    return ANUGA_OK;
}

static void timer_begin(anuga_domain *d, int t)
{
    if (!d->timing) return;
    KernelTimer &T = d->timers[t];
    if (T.start.empty()) {
        T.start.resize(kTimerRing);
        T.stop.resize(kTimerRing);
        for (int i = 0; i < kTimerRing; i++) {
            hipEventCreate(&T.start[i]);
            hipEventCreate(&T.stop[i]);
        }
    }
    if (T.pending == kTimerRing) {
        hipEventSynchronize(T.stop[T.pending - 1]);
        for (int i = 0; i < T.pending; i++) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, T.start[i], T.stop[i]);
            T.total_ms += ms;
        }
        T.launches += T.pending;
        T.pending = 0;
    }
    hipEventRecord(T.start[T.pending], d->stream);
}

static void timer_end(anuga_domain *d, int t)
{
    if (!d->timing) return;
    KernelTimer &T = d->timers[t];
    hipEventRecord(T.stop[T.pending], d->stream);
    T.pending++;
}

static void timer_flush(anuga_domain *d)
{
    for (int t = 0; t < T_COUNT; t++) {
        KernelTimer &T = d->timers[t];
        if (!T.pending) continue;
        hipEventSynchronize(T.stop[T.pending - 1]);
        for (int i = 0; i < T.pending; i++) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, T.start[i], T.stop[i]);
            T.total_ms += ms;
        }
        T.launches += T.pending;
        T.pending = 0;
    }
}


// ---------------------------------------------------------------------------
// kernel launches
// ---------------------------------------------------------------------------
static DeMesh dev_mesh(const anuga_domain *d)
{
    DeMesh m;
    m.N = d->N;
    m.nb = d->nb;
    m.cen = d->cen;
    m.gx = d->gx;
    m.gf = d->gf;
    m.area = d->area;
    m.nx = d->nx;
    m.ny = d->ny;
    m.len = d->len;
    m.radius = d->radius;
    return m;
}

static int launch_extrapolate(anuga_domain *d, DeQ *S)
{
    timer_begin(d, T_EXTRAP);
    if (d->mixed)
        hipLaunchKernelGGL(de_extrapolate<float>, dim3(d->nblocks), dim3(DE_BLOCK), 0, d->stream,
                           dev_mesh(d), dev_params(d->P), S, d->E, d->flags,
                           d->part + 2 * (size_t)d->nblocks, d->ctl);
    else
        hipLaunchKernelGGL(de_extrapolate<double>, dim3(d->nblocks), dim3(DE_BLOCK), 0, d->stream,
                           dev_mesh(d), dev_params(d->P), S, d->E, d->flags,
                           d->part + 2 * (size_t)d->nblocks, d->ctl);
    timer_end(d, T_EXTRAP);
    return ANUGA_OK;
}

static void launch_materialize(anuga_domain *d, DeQ *S)
{
    hipLaunchKernelGGL(de_materialize, dim3(d->nblocks), dim3(DE_BLOCK), 0, d->stream, d->N, S,
                       d->flags);
}

template <class T, int SUB>
static void launch_flux_t(anuga_domain *d, DeQ *S, DeQ *bk)
{
    DeBoundary bc{d->bkind, d->bval, d->M};
    if constexpr (sizeof(T) == 8)
        hipLaunchKernelGGL((de_flux<double, SUB>), dim3(d->nblocks), dim3(DE_BLOCK), 0, d->stream,
                           dev_mesh(d), dev_params(d->P), S, d->eta,
                           (const double *)d->E, bc, d->flags, d->ex[0], d->ex[1],
                           d->ex[2], bk, d->ctl, d->part, d->part + d->nblocks, d->dneg);
    else
        hipLaunchKernelGGL((de_flux_mixed<SUB>), dim3(d->nblocks), dim3(DE_BLOCK), 0, d->stream,
                           dev_mesh(d), dev_params(d->P), S, d->eta, (const DeEdgeRec32 *)d->E,
                           bc, d->flags, d->ex[0], d->ex[1], d->ex[2], bk, d->ctl, d->part,
                           d->part + d->nblocks, d->dneg);
}

// substep 0: explicit updates; substep 1: advance S with ctl->dt into the backup `bk`
static void launch_flux(anuga_domain *d, int substep, DeQ *S, DeQ *bk)
{
    timer_begin(d, substep ? T_FLUX1 : T_FLUX0);
    if (d->mixed) {
        if (substep) launch_flux_t<float, 1>(d, S, bk);
        else launch_flux_t<float, 0>(d, S, bk);
    } else {
        if (substep) launch_flux_t<double, 1>(d, S, bk);
        else launch_flux_t<double, 0>(d, S, bk);
    }
    timer_end(d, substep ? T_FLUX1 : T_FLUX0);
}

static void launch_update(anuga_domain *d, DeQ *S, DeQ *out)
{
    timer_begin(d, T_UPDATE);
    if (d->mixed)
        hipLaunchKernelGGL(de_update<false>, dim3(d->nblocks), dim3(DE_BLOCK), 0, d->stream, d->N,
                           dev_params(d->P), S, d->eta, d->flags, d->ex[0], d->ex[1], d->ex[2],
                           d->ctl, out, d->dneg);
    else
        hipLaunchKernelGGL(de_update<true>, dim3(d->nblocks), dim3(DE_BLOCK), 0, d->stream, d->N,
                           dev_params(d->P), S, d->eta, d->flags, d->ex[0], d->ex[1], d->ex[2],
                           d->ctl, out, d->dneg);
    timer_end(d, T_UPDATE);
}

static void launch_reduce(anuga_domain *d, int mode)
{
    const double *pdt = mode == 0 ? d->part : nullptr;
    const double *pbf = mode <= 1 ? d->part + d->nblocks : nullptr;
    const double *pm = (mode == 0 || mode >= 2) ? d->part + 2 * (size_t)d->nblocks : nullptr;
    hipLaunchKernelGGL(de_reduce, dim3(1), dim3(1024), 0, d->stream, d->nblocks, pdt, pbf, pm,
                       d->ctl, mode);
}

static void launch_single(anuga_domain *d, void (*k)(DeControl *))
{
    hipLaunchKernelGGL(k, dim3(1), dim3(64), 0, d->stream, d->ctl);
}

static void launch_volume(anuga_domain *d, DeQ *S)
{
    hipLaunchKernelGGL(de_volume, dim3(d->nblocks), dim3(DE_BLOCK), 0, d->stream, d->N, S, d->area,
                       d->part + 2 * (size_t)d->nblocks);
}

// leave device-controlled mode before host-controlled launches
static void ensure_live(anuga_domain *d)
{
    if (d->done_on_device) {
        launch_single(d, de_clear_done);
        d->done_on_device = false;
    }
}

// One device-controlled step: both substeps, the timestep on the device and
// the end-of-step bookkeeping. Every kernel returns at once when ctl->done.
static void enqueue_device_step(anuga_domain *d)
{
    const bool rk2 = d->P.timestep_fluxcalls == 2;
    launch_extrapolate(d, d->A);
    launch_flux(d, 0, d->A, nullptr);
    launch_reduce(d, 0);
    launch_single(d, de_timestep);
    launch_update(d, d->A, rk2 ? d->B : d->A);
    if (rk2) {
        launch_extrapolate(d, d->B);
        launch_flux(d, 1, d->B, d->A);
        launch_reduce(d, 1);
    }
    hipLaunchKernelGGL(de_finish_step, dim3(1), dim3(64), 0, d->stream, d->ctl, d->hist);
}

static void drop_graph(anuga_domain *d)
{
    if (d->graph_exec) hipGraphExecDestroy(d->graph_exec);
    if (d->graph) hipGraphDestroy(d->graph);
    d->graph_exec = nullptr;
    d->graph = nullptr;
}

// read the control block and counters; raise a pending device error
static int sync_scalars(anuga_domain *d)
{
    HIPCHK(d, hipMemcpyAsync(d->hctl, d->ctl, sizeof(DeControl), hipMemcpyDeviceToHost, d->stream));
    HIPCHK(d, hipMemcpyAsync(d->hneg, d->dneg, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             d->stream));
    HIPCHK(d, hipStreamSynchronize(d->stream));
    const int e = d->hctl->status;
    if (e) {
        hipMemsetAsync(&d->ctl->status, 0, sizeof(int), d->stream);
        hipStreamSynchronize(d->stream);
        d->hctl->status = 0;
        if (e & DE_ERR_SEMI_IMPLICIT)
            return fail(d, ANUGA_E_SEMI_IMPLICIT, "update: division by zero in semi implicit update");
        if (e & DE_ERR_NO_NEIGHBOUR) return fail(d, ANUGA_E_NO_NEIGHBOUR, "Internal neighbour not found");
        if (e & DE_ERR_SMALL_TIMESTEP)
            return fail(d, ANUGA_E_TIMESTEP, "WARNING: Too small timestep reached even after "
                                             "max_smallsteps steps of 1 order scheme");
        if (e & DE_ERR_OVERSHOOT)
            return fail(d, ANUGA_E_TIMESTEP, "WARNING (domain.py): time overshot finaltime. ");
    }
    return ANUGA_OK;
}

static int launch_check(anuga_domain *d)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(d, ANUGA_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return ANUGA_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *anuga_last_error(const anuga_domain *d)
{
    return d ? d->err.c_str() : g_create_error.c_str();
}

int anuga_create(const anuga_mesh *mesh, const anuga_params *params, anuga_domain **out)
{
    if (!out) return fail(nullptr, ANUGA_E_ARG, "out is NULL");
    *out = nullptr;
    if (!mesh) return fail(nullptr, ANUGA_E_ARG, "mesh is NULL");
    int rc = check_params(nullptr, params);
    if (rc) return rc;
    const int64_t N64 = mesh->number_of_elements, M64 = mesh->boundary_length;
    if (N64 <= 0 || N64 >= (int64_t(1) << 29))
        return fail(nullptr, ANUGA_E_ARG, "number_of_elements must be in [1, 2^29)");
    if (M64 < 0 || M64 >= (int64_t(1) << 30)) return fail(nullptr, ANUGA_E_ARG, "bad boundary_length");
    const int N = (int)N64, M = (int)M64;
    const int64_t *nbr = mesh->neighbours, *nbe = mesh->neighbour_edges;
    const int64_t *sur = mesh->surrogate_neighbours, *nob = mesh->number_of_boundaries;
    if (!nbr || !nbe || !sur || !nob || !mesh->normals || !mesh->edgelengths || !mesh->radii ||
        !mesh->areas || !mesh->centroid_coordinates || !mesh->edge_coordinates)
        return fail(nullptr, ANUGA_E_ARG, "a mesh array is NULL");

    // validate connectivity: the device derives surrogate neighbours and the
    // boundary count from `neighbours`, which requires no internal boundaries
    for (int k = 0; k < N; k++) {
        int nb = 0;
        for (int i = 0; i < 3; i++) {
            const int64_t ki = 3 * (int64_t)k + i, n = nbr[ki];
            if (n >= 0) {
                if (n >= N) return fail(nullptr, ANUGA_E_ARG, "neighbour index out of range");
                const int64_t m = nbe[ki];
                if (m < 0 || m > 2 || nbr[3 * n + m] != k)
                    return fail(nullptr, ANUGA_E_ARG, "inconsistent neighbour_edges");
                if (sur[ki] != n)
                    return fail(nullptr, ANUGA_E_ARG, "internal boundary edges are not supported");
            } else {
                if (-n - 1 >= M) return fail(nullptr, ANUGA_E_ARG, "boundary index out of range");
                if (sur[ki] != k)
                    return fail(nullptr, ANUGA_E_ARG, "internal boundary edges are not supported");
                nb++;
            }
        }
        if (nob[k] != nb) return fail(nullptr, ANUGA_E_ARG, "number_of_boundaries inconsistent");
    }

    anuga_domain *d = new anuga_domain();
    d->N = N;
    d->M = M;
    d->P = *params;
    d->nblocks = (N + DE_BLOCK - 1) / DE_BLOCK;

    d->perm.resize(N);
    d->iperm.resize(N);
    if (params->reorder) {
        d->perm = rcm_order(N, nbr);
    } else {
        for (int k = 0; k < N; k++) d->perm[k] = k;
    }
    for (int p = 0; p < N; p++) d->iperm[d->perm[p]] = p;

#define DCHK(call)                                                                       \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            std::string m_ = std::string(#call) + ": " + hipGetErrorString(e_);          \
            anuga_destroy(d);                                                            \
            return fail(nullptr, ANUGA_E_HIP, m_);                                       \
        }                                                                                \
    } while (0)

    DCHK(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
    d->mixed = params->precision == ANUGA_PRECISION_MIXED;
    const size_t n3 = 3 * (size_t)N;
    DCHK(dalloc(&d->nb, N));
    DCHK(dalloc(&d->cen, N));
    DCHK(dalloc(&d->area, N));
    DCHK(dalloc(&d->eta, N));
    DCHK(dalloc(&d->A, N));
    DCHK(dalloc(&d->B, N));
    for (int q = 0; q < 3; q++) DCHK(dalloc(&d->ex[q], N));
    if (d->mixed) {
        DCHK(dalloc((DeGeoX32 **)&d->gx, N));
        DCHK(dalloc((DeGeoF32 **)&d->gf, N));
        DCHK(dalloc((DeEdgeRec32 **)&d->E, N));
    } else {
        DCHK(dalloc((DeGeoX64 **)&d->gx, N));
        DCHK(dalloc(&d->nx, n3));
        DCHK(dalloc(&d->ny, n3));
        DCHK(dalloc(&d->len, n3));
        DCHK(dalloc(&d->radius, N));
        DCHK(dalloc((double **)&d->E, 4 * n3));
    }
    DCHK(dalloc(&d->flags, N));
    DCHK(dalloc(&d->bkind, M));
    DCHK(dalloc(&d->bval, 3 * (size_t)M));
    DCHK(dalloc(&d->part, 3 * (size_t)d->nblocks));
    DCHK(dalloc(&d->ctl, 1));
    DCHK(dalloc(&d->dneg, 1));
    DCHK(hipHostMalloc((void **)&d->hctl, sizeof(DeControl), hipHostMallocDefault));
    DCHK(hipHostMalloc((void **)&d->hneg, sizeof(unsigned long long), hipHostMallocDefault));

    {
        const double *nrm = mesh->normals, *el = mesh->edgelengths;
        const double *cen = mesh->centroid_coordinates, *em = mesh->edge_coordinates;
        std::vector<int4> hn(N);
        std::vector<double2> hc(N);
        std::vector<double> ha(N);
        for (int p = 0; p < N; p++) {
            const int64_t k = d->perm[p];
            int v[3];
            for (int i = 0; i < 3; i++) {
                const int64_t n = nbr[3 * k + i];
                v[i] = n >= 0 ? ((d->iperm[n] << 2) | (int)nbe[3 * k + i]) : (int)n;
            }
            hn[p] = make_int4(v[0], v[1], v[2], (int)nob[k]);
            hc[p] = make_double2(cen[2 * k], cen[2 * k + 1]);
            ha[p] = mesh->areas[k];
        }
        DCHK(hipMemcpy(d->nb, hn.data(), N * sizeof(int4), hipMemcpyHostToDevice));
        DCHK(hipMemcpy(d->cen, hc.data(), N * sizeof(double2), hipMemcpyHostToDevice));
        DCHK(hipMemcpy(d->area, ha.data(), N * sizeof(double), hipMemcpyHostToDevice));
        if (!d->mixed) {
            std::vector<DeGeoX64> gx(N);
            std::vector<double> a(n3), b(n3), c(n3), r(N);
            for (int p = 0; p < N; p++) {
                const int64_t k = d->perm[p];
                for (int i = 0; i < 3; i++) {
                    // xv - x and yv - y exactly as sw_domain_orig.c:2130-2135 forms them
                    gx[p].dxv[i] = em[6 * k + 2 * i] - cen[2 * k];
                    gx[p].dyv[i] = em[6 * k + 2 * i + 1] - cen[2 * k + 1];
                    a[(size_t)i * N + p] = nrm[6 * k + 2 * i];
                    b[(size_t)i * N + p] = nrm[6 * k + 2 * i + 1];
                    c[(size_t)i * N + p] = el[3 * k + i];
                }
                r[p] = mesh->radii[k];
            }
            DCHK(hipMemcpy(d->gx, gx.data(), N * sizeof(DeGeoX64), hipMemcpyHostToDevice));
            DCHK(hipMemcpy(d->nx, a.data(), n3 * sizeof(double), hipMemcpyHostToDevice));
            DCHK(hipMemcpy(d->ny, b.data(), n3 * sizeof(double), hipMemcpyHostToDevice));
            DCHK(hipMemcpy(d->len, c.data(), n3 * sizeof(double), hipMemcpyHostToDevice));
            DCHK(hipMemcpy(d->radius, r.data(), N * sizeof(double), hipMemcpyHostToDevice));
        } else {
            std::vector<DeGeoX32> gx(N);
            std::vector<DeGeoF32> gf(N);
            for (int p = 0; p < N; p++) {
                const int64_t k = d->perm[p];
                for (int i = 0; i < 3; i++) {
                    gx[p].dxv[i] = (float)(em[6 * k + 2 * i] - cen[2 * k]);
                    gx[p].dyv[i] = (float)(em[6 * k + 2 * i + 1] - cen[2 * k + 1]);
                    const int64_t s = sur[3 * k + i];
                    gx[p].rx[i] = (float)(cen[2 * s] - cen[2 * k]);
                    gx[p].ry[i] = (float)(cen[2 * s + 1] - cen[2 * k + 1]);
                    gf[p].nx[i] = (float)nrm[6 * k + 2 * i];
                    gf[p].ny[i] = (float)nrm[6 * k + 2 * i + 1];
                    gf[p].len[i] = (float)el[3 * k + i];
                }
                gf[p].radius = (float)mesh->radii[k];
            }
            for (int p = 0; p < N; p++)
                for (int i = 0; i < 3; i++) gf[p].dzc[i] = gf[p].pad[i] = 0.f;
            DCHK(hipMemcpy(d->gx, gx.data(), N * sizeof(DeGeoX32), hipMemcpyHostToDevice));
            DCHK(hipMemcpy(d->gf, gf.data(), N * sizeof(DeGeoF32), hipMemcpyHostToDevice));
            d->gf32 = std::move(gf);
        }
    }
    DCHK(hipMemset(d->eta, 0, N * sizeof(double)));
    DCHK(hipMemset(d->A, 0, N * sizeof(DeQ)));
    DCHK(hipMemset(d->B, 0, N * sizeof(DeQ)));
    for (int q = 0; q < 3; q++) DCHK(hipMemset(d->ex[q], 0, N * sizeof(double)));
    DCHK(hipMemset(d->E, 0, d->mixed ? N * sizeof(DeEdgeRec32) : 4 * n3 * sizeof(double)));
    DCHK(hipMemset(d->flags, DE_F_KEEP, N));
    DCHK(hipMemset(d->bkind, 0, std::max(M, 1)));
    DCHK(hipMemset(d->bval, 0, 3 * (size_t)std::max(M, 1) * sizeof(double)));
    DCHK(hipMemset(d->ctl, 0, sizeof(DeControl)));
    DCHK(hipMemset(d->dneg, 0, sizeof(unsigned long long)));
    std::memset(d->hctl, 0, sizeof(DeControl));
    *d->hneg = 0;
    {
        const char *g = std::getenv("ANUGA_HIP_GRAPH");
        d->use_graph = !(g && g[0] == '0');
    }
    DCHK(hipDeviceSynchronize());
#undef DCHK
    *out = d;
    return ANUGA_OK;
}

void anuga_destroy(anuga_domain *d)
{
    if (!d) return;
    if (d->stream) hipStreamSynchronize(d->stream);
    drop_graph(d);
    void *ptrs[] = {d->nb, d->cen, d->gx, d->gf, d->area, d->nx, d->ny, d->len, d->radius,
                    d->eta, d->A, d->B, d->E, d->flags,
                    d->ex[0], d->ex[1], d->ex[2], d->bkind, d->bval, d->part, d->ctl, d->dneg,
                    d->hist};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (d->hctl) hipHostFree(d->hctl);
    if (d->hneg) hipHostFree(d->hneg);
    for (int t = 0; t < T_COUNT; t++) {
        for (auto e : d->timers[t].start) hipEventDestroy(e);
        for (auto e : d->timers[t].stop) hipEventDestroy(e);
    }
    if (d->stream) hipStreamDestroy(d->stream);
    delete d;
}

int anuga_set_params(anuga_domain *d, const anuga_params *params)
{
    if (!d) return ANUGA_E_ARG;
    int rc = check_params(d, params);
    if (rc) return rc;
    if (params->precision != d->P.precision || params->reorder != d->P.reorder)
        return fail(d, ANUGA_E_ARG, "precision and reorder are fixed at creation");
    if (params->timestep_fluxcalls != d->P.timestep_fluxcalls && d->substate != 0)
        return fail(d, ANUGA_E_STATE, "cannot change the timestepping method inside a step");
    if (std::memcmp(params, &d->P, sizeof(anuga_params)) != 0) drop_graph(d);
    d->P = *params;
    return ANUGA_OK;
}

// component of the DeQ record holding a quantity (-1: not in the record)
static int record_slot(int q)
{
    switch (q) {
    case ANUGA_STAGE: return 0;
    case ANUGA_XMOMENTUM: return 1;
    case ANUGA_YMOMENTUM: return 2;
    case ANUGA_ELEVATION: return 3;
    default: return -1;
    }
}

static int download_records(anuga_domain *d, DeQ *src, std::vector<DeQ> &h)
{
    h.resize(d->N);
    HIPCHK(d, hipStreamSynchronize(d->stream));
    HIPCHK(d, hipMemcpy(h.data(), src, d->N * sizeof(DeQ), hipMemcpyDeviceToHost));
    return ANUGA_OK;
}

int anuga_set_centroid_values(anuga_domain *d, int quantity, const double *values)
{
    if (!d || !values) return ANUGA_E_ARG;
    const int N = d->N;
    if (quantity == ANUGA_FRICTION) {
        std::vector<double> t(N);
        for (int p = 0; p < N; p++) t[p] = values[d->perm[p]];
        HIPCHK(d, hipStreamSynchronize(d->stream));
        HIPCHK(d, hipMemcpy(d->eta, t.data(), N * sizeof(double), hipMemcpyHostToDevice));
        return ANUGA_OK;
    }
    const int slot = record_slot(quantity);
    if (slot < 0) return fail(d, ANUGA_E_ARG, "quantity cannot be set");
    // the bed is part of both state buffers; the others of the current one
    DeQ *bufs[2] = {d->substate ? d->B : d->A, slot == 3 ? (d->substate ? d->A : d->B) : nullptr};
    for (DeQ *buf : bufs) {
        if (!buf) continue;
        std::vector<DeQ> h;
        int rc = download_records(d, buf, h);
        if (rc) return rc;
        for (int p = 0; p < N; p++) (&h[p].w)[slot] = values[d->perm[p]];
        HIPCHK(d, hipMemcpy(buf, h.data(), N * sizeof(DeQ), hipMemcpyHostToDevice));
    }
    if (slot == 3 && d->mixed) {
        // bed offsets of the mixed flux's local datum
        std::vector<int4> nb(N);
        HIPCHK(d, hipMemcpy(nb.data(), d->nb, N * sizeof(int4), hipMemcpyDeviceToHost));
        for (int p = 0; p < N; p++) {
            const int v[3] = {nb[p].x, nb[p].y, nb[p].z};
            const double zp = values[d->perm[p]];
            for (int i = 0; i < 3; i++)
                d->gf32[p].dzc[i] = v[i] >= 0 ? (float)(values[d->perm[v[i] >> 2]] - zp) : 0.f;
        }
        HIPCHK(d, hipMemcpy(d->gf, d->gf32.data(), N * sizeof(DeGeoF32), hipMemcpyHostToDevice));
    }
    return ANUGA_OK;
}

int anuga_get_centroid_values(anuga_domain *d, int quantity, double *values)
{
    if (!d || !values) return ANUGA_E_ARG;
    const int N = d->N;
    if (quantity == ANUGA_FRICTION) {
        std::vector<double> t(N);
        HIPCHK(d, hipStreamSynchronize(d->stream));
        HIPCHK(d, hipMemcpy(t.data(), d->eta, N * sizeof(double), hipMemcpyDeviceToHost));
        for (int p = 0; p < N; p++) values[d->perm[p]] = t[p];
        return ANUGA_OK;
    }
    const int slot = record_slot(quantity);
    if (slot < 0 && quantity != ANUGA_HEIGHT) return fail(d, ANUGA_E_ARG, "unknown quantity");
    std::vector<DeQ> h;
    int rc = download_records(d, d->substate ? d->B : d->A, h);
    if (rc) return rc;
    for (int p = 0; p < N; p++)
        values[d->perm[p]] = slot >= 0 ? (&h[p].w)[slot] : std::fmax(h[p].w - h[p].z, 0.0);
    return ANUGA_OK;
}

int anuga_get_edge_values(anuga_domain *d, int quantity, double *values)
{
    if (!d || !values) return ANUGA_E_ARG;
    const size_t n3 = 3 * (size_t)d->N;
    if (quantity != ANUGA_STAGE && quantity != ANUGA_HEIGHT && quantity != ANUGA_XMOMENTUM &&
        quantity != ANUGA_YMOMENTUM && quantity != ANUGA_ELEVATION)
        return fail(d, ANUGA_E_ARG, "unknown edge quantity");
    HIPCHK(d, hipStreamSynchronize(d->stream));
    std::vector<double> e(4 * n3);
    if (d->mixed) {
        // edge = centroid + jump; the centroid state is the one the last
        // extrapolation saw after it was written back (distribute)
        std::vector<DeEdgeRec32> r(d->N);
        std::vector<DeQ> st;
        HIPCHK(d, hipMemcpy(r.data(), d->E, d->N * sizeof(DeEdgeRec32), hipMemcpyDeviceToHost));
        int rc = download_records(d, d->substate ? d->B : d->A, st);
        if (rc) return rc;
        for (int p = 0; p < d->N; p++)
            for (int i = 0; i < 3; i++) {
                const float4 f = r[p].e[i];
                const double hc = r[p].hc;
                const double wc = st[p].z + hc;
                double *o = &e[4 * (3 * (size_t)p + i)];
                o[0] = wc + f.x;
                o[1] = hc + f.y;
                o[2] = f.z;
                o[3] = f.w;
            }
    } else {
        // four [edge][triangle] arrays -> (triangle, edge, quantity)
        std::vector<double> soa(4 * n3);
        HIPCHK(d, hipMemcpy(soa.data(), d->E, 4 * n3 * sizeof(double), hipMemcpyDeviceToHost));
        for (int p = 0; p < d->N; p++)
            for (int i = 0; i < 3; i++)
                for (int qq = 0; qq < 4; qq++)
                    e[4 * (3 * (size_t)p + i) + qq] = soa[qq * n3 + (size_t)i * d->N + p];
    }
    const int N = d->N;
    for (int p = 0; p < N; p++) {
        const int64_t k = d->perm[p];
        for (int i = 0; i < 3; i++) {
            const double *r = &e[4 * (3 * (size_t)p + i)];   // w, h, uh, vh
            double v;
            switch (quantity) {
            case ANUGA_STAGE: v = r[0]; break;
            case ANUGA_HEIGHT: v = r[1]; break;
            case ANUGA_XMOMENTUM: v = r[2]; break;
            case ANUGA_YMOMENTUM: v = r[3]; break;
            default: v = r[0] - r[1]; break;  // bed_edge = stage_edge - height_edge
            }
            values[3 * k + i] = v;
        }
    }
    return ANUGA_OK;
}

int anuga_set_boundary(anuga_domain *d, const int32_t *kind, const double *dirichlet_values)
{
    if (!d) return ANUGA_E_ARG;
    const int M = d->M;
    if (M > 0 && !kind) return fail(d, ANUGA_E_ARG, "kind is NULL");
    std::vector<signed char> k(std::max(M, 1), 0);
    std::vector<double> v(3 * (size_t)std::max(M, 1), 0.0);
    for (int j = 0; j < M; j++) {
        if (kind[j] < 0 || kind[j] > 3) return fail(d, ANUGA_E_ARG, "unknown boundary kind");
        k[j] = (signed char)kind[j];
        if (kind[j] == ANUGA_BC_DIRICHLET) {
            if (!dirichlet_values) return fail(d, ANUGA_E_ARG, "Dirichlet values are NULL");
            for (int q = 0; q < 3; q++) v[(size_t)q * M + j] = dirichlet_values[3 * (size_t)j + q];
        }
    }
    HIPCHK(d, hipStreamSynchronize(d->stream));
    HIPCHK(d, hipMemcpy(d->bkind, k.data(), k.size(), hipMemcpyHostToDevice));
    HIPCHK(d, hipMemcpy(d->bval, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
    d->boundary_set = true;
    return ANUGA_OK;
}

int anuga_distribute_to_vertices_and_edges(anuga_domain *d, double *mass_error)
{
    if (!d) return ANUGA_E_ARG;
    ensure_live(d);
    DeQ *S = d->substate ? d->B : d->A;
    int rc = launch_extrapolate(d, S);
    if (rc) return rc;
    launch_materialize(d, S);
    launch_reduce(d, 2);
    rc = launch_check(d);
    if (rc) return rc;
    rc = sync_scalars(d);
    if (rc) return rc;
    if (mass_error) *mass_error = d->hctl->mass;
    return ANUGA_OK;
}

int anuga_compute_fluxes(anuga_domain *d, double *flux_timestep)
{
    if (!d) return ANUGA_E_ARG;
    if (d->substate != 0) return fail(d, ANUGA_E_STATE, "compute_fluxes called inside an rk2 step");
    if (!d->boundary_set && d->M > 0) return fail(d, ANUGA_E_STATE, "boundary conditions not set");
    ensure_live(d);
    int rc = launch_extrapolate(d, d->A);
    if (rc) return rc;
    launch_flux(d, 0, d->A, nullptr);
    launch_reduce(d, 0);
    rc = launch_check(d);
    if (rc) return rc;
    rc = sync_scalars(d);
    if (rc) return rc;
    if (d->timing) timer_flush(d);
    if (flux_timestep) *flux_timestep = d->hctl->flux_dt;
    return ANUGA_OK;
}

static int read_negative(anuga_domain *d, int64_t *num_negative)
{
    int rc = sync_scalars(d);
    if (rc) return rc;
    const unsigned long long n = *d->hneg;
    *num_negative = (int64_t)(n - d->neg_seen);
    d->neg_seen = n;
    return ANUGA_OK;
}

int anuga_update_conserved_quantities(anuga_domain *d, double dt, int64_t *num_negative)
{
    if (!d) return ANUGA_E_ARG;
    if (d->substate != 0) return fail(d, ANUGA_E_STATE, "update called twice in a step");
    ensure_live(d);
    const bool rk2 = d->P.timestep_fluxcalls == 2;
    hipLaunchKernelGGL(de_set_dt, dim3(1), dim3(64), 0, d->stream, d->ctl, dt);
    launch_update(d, d->A, rk2 ? d->B : d->A);
    if (!rk2) hipLaunchKernelGGL(de_integrate, dim3(1), dim3(64), 0, d->stream, d->ctl, 0);
    int rc = launch_check(d);
    if (rc) return rc;
    if (rk2) d->substate = 1;
    return num_negative ? read_negative(d, num_negative) : ANUGA_OK;
}

int anuga_rk2_second_substep(anuga_domain *d, double dt, int64_t *num_negative)
{
    if (!d) return ANUGA_E_ARG;
    if (d->substate != 1) return fail(d, ANUGA_E_STATE, "second substep without a first");
    hipLaunchKernelGGL(de_set_dt, dim3(1), dim3(64), 0, d->stream, d->ctl, dt);
    int rc = launch_extrapolate(d, d->B);
    if (rc) return rc;
    launch_flux(d, 1, d->B, d->A);
    launch_reduce(d, 1);
    hipLaunchKernelGGL(de_integrate, dim3(1), dim3(64), 0, d->stream, d->ctl, 1);
    rc = launch_check(d);
    if (rc) return rc;
    d->substate = 0;
    return num_negative ? read_negative(d, num_negative) : ANUGA_OK;
}

int anuga_evolve_steps(anuga_domain *d, int64_t max_steps, const anuga_time_control *tc,
                       anuga_step_state *st, double *dt_history)
{
    if (!d || !tc || !st || max_steps < 0) return ANUGA_E_ARG;
    if (d->substate != 0) return fail(d, ANUGA_E_STATE, "evolve_steps called inside an rk2 step");
    if (!d->boundary_set && d->M > 0) return fail(d, ANUGA_E_STATE, "boundary conditions not set");
    if (dt_history && max_steps > d->hist_cap) {
        if (d->hist) hipFree(d->hist);
        d->hist = nullptr;
        d->hist_cap = 0;
        drop_graph(d);
        HIPCHK(d, dalloc(&d->hist, (size_t)max_steps));
        d->hist_cap = max_steps;
    }
    DeBatchSetup b;
    b.t = st->relative_time;
    b.rec_min = st->recorded_min_timestep;
    b.rec_max = st->recorded_max_timestep;
    b.cfl = tc->CFL;
    b.max_dt = tc->evolve_max_timestep;
    b.min_dt = tc->evolve_min_timestep;
    b.finaltime = tc->relative_finaltime;
    b.yieldtime = tc->relative_yieldtime;
    b.fixed_flux_dt = tc->fixed_flux_timestep;
    b.eps = tc->epsilon;
    b.smallsteps = (int)st->smallsteps;
    b.order = (int)st->order;
    b.default_order = (int)tc->default_order;
    b.max_smallsteps = (int)tc->max_smallsteps;
    b.hist_cap = dt_history ? (int)std::min<int64_t>(max_steps, d->hist_cap) : 0;
    b.rk2 = d->P.timestep_fluxcalls == 2;
    hipLaunchKernelGGL(de_setup_batch, dim3(1), dim3(64), 0, d->stream, d->ctl, b);
    d->done_on_device = true;

    const bool graph = d->use_graph && !d->timing;
    if (graph && !d->graph_exec && max_steps > 0) {
        HIPCHK(d, hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal));
        enqueue_device_step(d);
        hipGraph_t g = nullptr;
        HIPCHK(d, hipStreamEndCapture(d->stream, &g));
        d->graph = g;
        HIPCHK(d, hipGraphInstantiate(&d->graph_exec, d->graph, nullptr, nullptr, 0));
    }
    for (int64_t i = 0; i < max_steps; i++) {
        if (graph) HIPCHK(d, hipGraphLaunch(d->graph_exec, d->stream));
        else enqueue_device_step(d);
    }
    int rc = launch_check(d);
    if (rc) return rc;
    rc = sync_scalars(d);
    const DeControl &c = *d->hctl;
    st->relative_time = c.t;
    st->timestep = c.dt;
    st->flux_timestep = c.flux_dt;
    st->recorded_min_timestep = c.rec_min;
    st->recorded_max_timestep = c.rec_max;
    st->smallsteps = c.smallsteps;
    st->order = c.order;
    st->steps_taken = c.nsteps;
    st->reached = c.done;
    if (dt_history && c.nsteps > 0) {
        HIPCHK(d, hipMemcpy(dt_history, d->hist,
                            (size_t)std::min<int64_t>(c.nsteps, max_steps) * sizeof(double),
                            hipMemcpyDeviceToHost));
    }
    return rc;
}

int anuga_get_boundary_flux_sums(anuga_domain *d, double *sums)
{
    if (!d || !sums) return ANUGA_E_ARG;
    int rc = sync_scalars(d);
    if (rc) return rc;
    sums[0] = d->hctl->bflux0;
    sums[1] = d->P.timestep_fluxcalls == 2 ? d->hctl->bflux1 : 0.0;
    return ANUGA_OK;
}

int anuga_get_boundary_flux_integral(anuga_domain *d, double *integral)
{
    if (!d || !integral) return ANUGA_E_ARG;
    int rc = sync_scalars(d);
    if (rc) return rc;
    *integral = d->hctl->integral;
    return ANUGA_OK;
}

int anuga_get_negative_cells(anuga_domain *d, int64_t *total)
{
    if (!d || !total) return ANUGA_E_ARG;
    *total = (int64_t)*d->hneg;
    return ANUGA_OK;
}

int anuga_get_water_volume(anuga_domain *d, double *volume)
{
    if (!d || !volume) return ANUGA_E_ARG;
    DeQ *S = d->substate ? d->B : d->A;
    launch_volume(d, S);
    launch_reduce(d, 3);
    int rc = launch_check(d);
    if (rc) return rc;
    rc = sync_scalars(d);
    if (rc) return rc;
    *volume = d->hctl->volume;
    return ANUGA_OK;
}

int anuga_rcm_order(int64_t N, const int64_t *neighbours, int64_t *order)
{
    if (N <= 0 || N >= (int64_t(1) << 29) || !neighbours || !order) return ANUGA_E_ARG;
    for (int64_t i = 0; i < 3 * N; i++)
        if (neighbours[i] >= N) return ANUGA_E_ARG;
    const std::vector<int> o = rcm_order((int)N, neighbours);
    for (int64_t i = 0; i < N; i++) order[i] = o[i];
    return ANUGA_OK;
}

int anuga_synchronize(anuga_domain *d)
{
    if (!d) return ANUGA_E_ARG;
    HIPCHK(d, hipStreamSynchronize(d->stream));
    return ANUGA_OK;
}

int anuga_reset_kernel_timing(anuga_domain *d, int enable)
{
    if (!d) return ANUGA_E_ARG;
    HIPCHK(d, hipStreamSynchronize(d->stream));
    for (int t = 0; t < T_COUNT; t++) {
        d->timers[t].pending = 0;
        d->timers[t].total_ms = 0.0;
        d->timers[t].launches = 0;
    }
    d->timing = enable != 0;
    return ANUGA_OK;
}

int anuga_kernel_timing(anuga_domain *d, const char *name, double *avg_ms, int64_t *launches)
{
    if (!d || !name) return ANUGA_E_ARG;
    timer_flush(d);
    for (int t = 0; t < T_COUNT; t++) {
        if (std::strcmp(name, kTimerNames[t]) == 0) {
            const KernelTimer &T = d->timers[t];
            if (avg_ms) *avg_ms = T.launches ? T.total_ms / (double)T.launches : 0.0;
            if (launches) *launches = T.launches;
            return ANUGA_OK;
        }
    }
    return fail(d, ANUGA_E_ARG, std::string("unknown kernel timer ") + name);
}

}  // extern "C"

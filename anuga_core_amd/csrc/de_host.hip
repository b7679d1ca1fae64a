// Host runtime and C ABI of the DE step (see include/anuga_hip.h).
//
// A handle owns all device buffers of one or more scenarios (independent
// meshes stepped together: an ensemble). Each scenario's triangles are
// renumbered once at creation (recursive coordinate bisection of the
// centroids, so that each workgroup's run of triangles is a compact tile for
// the fused kernels, or Reverse-Cuthill-McKee on the triangle adjacency) so that
// neighbour gathers stay local; the permutation is applied on upload and
// undone on download, so callers only ever see the reference's order. Scenarios sit back to back in one internal
// numbering, each starting on a workgroup boundary, so that every workgroup
// belongs to exactly one scenario and reads that scenario's control block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "anuga_hip.h"
#include "de_kernels.hip"
#include "de_adjoint.hip"

namespace {

thread_local std::string g_create_error;

// ---------------------------------------------------------------------------
// the device scratch of anuga_water_volume_noalloc: a control block, the
// reduction's chunk results and one partial per block of at most
// kNoAllocMaxN triangles. One copy per GPU, alive until the process ends;
// host code uses the address hipGetSymbolAddress gives on each device, never
// the array's name.
// ---------------------------------------------------------------------------
static const int64_t kNoAllocMaxN = int64_t(1) << 27;
static const int kNoAllocCtl = 32;                        // doubles for the DeControl
static const int kNoAllocChunks = 3 * DE_RCHUNKS;         // the first stage's results
static const int64_t kNoAllocDoubles =
    kNoAllocCtl + kNoAllocChunks + (kNoAllocMaxN + DE_BLOCK - 1) / DE_BLOCK;
static_assert(sizeof(DeControl) <= kNoAllocCtl * sizeof(double), "control block size");
static __device__ __align__(256) double scratch[kNoAllocDoubles];
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[256];                        // per device, from hipGetSymbolAddress

struct KernelTimer {
    std::vector<hipEvent_t> start, stop;
    int pending = 0;
    double total_ms = 0.0;
    int64_t launches = 0;
};

const char *kTimerNames[] = {"extrapolate", "flux0", "update", "flux1", "xflux0", "xflux1"};
enum { T_EXTRAP = 0, T_FLUX0, T_UPDATE, T_FLUX1, T_XFLUX0, T_XFLUX1, T_COUNT };
const int kTimerRing = 64;

}  // namespace

// a Rate_operator's device data
struct RateOp {
    int kind = 0;            // 0: Rate_operator, 1: Set_stage_operator (rate = stage values),
                             // 2: Inlet_operator (rate = areas, influx = its 4 statistics)
    int n = 0;
    int *idx = nullptr;      // [n] internal triangles
    double *rate = nullptr;  // [n] rates, or NULL: the scalar rate0
    double rate0 = 0.0, factor = 1.0;
    int careful = 0;         // some rate < 0
    double *influx = nullptr;
    // inlets: np.sum(areas), Q at the operator's time and a timestep later,
    // velocity mode (0 region, 1 given, 2 zero) and the given velocity
    double total_area = 0.0, Q1 = 0.0, Q2 = 0.0, u = 0.0, v = 0.0;
    int vmode = 0;
};

struct anuga_domain {
    int N = 0, M = 0, nblocks = 0;   // N: internal triangles incl. padding; M: boundary edges
    int S = 1;                       // scenarios
    anuga_params P{};
    hipStream_t stream = nullptr;
    // per scenario: first internal triangle, triangles, first boundary edge, boundary edges
    std::vector<int> kbeg, ntri, mbeg, nbnd;
    std::vector<char> mesh_set, bnd_set;
    int meshes_missing = 0, boundaries_missing = 0;
    std::vector<int> perm;   // perm[internal] = caller's index within its scenario (-1: padding)

    // static mesh records, internal order
    int4 *nb = nullptr;
    double2 *cen = nullptr;
    void *gx = nullptr, *gf = nullptr;   // DeGeoX64/F64 or DeGeoX32/F32
    double *area = nullptr;
    double *nx = nullptr, *ny = nullptr, *len = nullptr, *radius = nullptr;  // FP64 flux geometry
    int *scen = nullptr;                 // [nblocks] scenario of each block (S > 1)
    std::vector<int4> hnb;               // host copy of nb
    int *caller = nullptr;               // [N] caller's index within the scenario (-1: padding)
    double *vx = nullptr;                // [N][6] vertex coordinates (sloped friction), or NULL
    // compact tiles (reorder 2, one FP64 scenario): per-tile halo half-edges
    // and per-triangle halo slots for the fused extrapolation + flux kernel;
    // hcap == 0: no tiles
    int *tstart = nullptr, *thalo = nullptr;
    unsigned *tslot = nullptr;
    int hcap = 0;
    // per scenario, until every mesh is set: tile offsets, halo (internal
    // ids), slots and largest halo (-1: no tiles)
    std::vector<std::vector<int>> ts_start, ts_halo, ts_estart, ts_esplit;
    std::vector<std::vector<unsigned>> ts_slot, ts_edge;
    // tile edge lists of the once-per-edge kernel (null: not built); mirror:
    // every interior edge's normals and lengths mirror exactly (checked per
    // scenario at upload), which that kernel assumes
    int *testart = nullptr, *tesplit = nullptr;
    unsigned *tedge = nullptr;
    bool mirror = true;
    bool edge_once = true;     // ANUGA_HIP_EDGE_ONCE=0: the gather-every-side kernel
    // vertex geometry: the fused FP64 kernel forms edge offsets, normals and
    // lengths from each triangle's vertex coordinates (vx) instead of reading
    // the stored records, where that reproduces every scenario's arrays bit
    // for bit (ANUGA_HIP_VERTEX_GEOMETRY=0: the stored records)
    bool vgeom = true;
    std::vector<int> ts_hcap;
    bool e_stale = false;                // E does not hold the last extrapolation
    int *rw = nullptr;                   // [3][N] riverwall record per half-edge, -1: none
    DeRiverwall *rwd = nullptr;          // riverwall records; rw == nullptr: no riverwalls
    // output (one scenario): node CSR of the centroid averaging, in the
    // reference's entry order, with internal triangle ids; caller -> internal map
    int nodes = 0;
    int *node_start = nullptr, *node_tri = nullptr, *iperm_dev = nullptr;
    float *out_f32 = nullptr;            // staging of one frame: [3][nodes] + [3][N]
    // gauges (one scenario): internal triangle ids, rows of times and states
    int G = 0, gcap = 0;
    int64_t grows_host = 0;              // rows on the device not yet downloaded
    int *gtri = nullptr;
    double *gtimes = nullptr, *gvals = nullptr;
    std::vector<RateOp> ops;             // fractional-step rate operators, in order
    std::vector<int> hiperm;             // caller's index -> internal (one scenario)
    int *sub_idx = nullptr;              // staging of a subset access
    DeQ *sub_val = nullptr;
    int sub_cap = 0;
    bool mixed = false;
    // quantities
    double *eta = nullptr;
    DeQ *A = nullptr;   // step state / rk2 backup: {w, uh, vh, z}
    DeQ *B = nullptr;   // rk2 substep state
    DeQ *C = nullptr;   // rk3 middle-substep state of the fused kernels (allocated on demand)
    void *E = nullptr;  // FP64: w,h,uh,vh [4][3][N]; mixed: DeEdgeRec32 [N]
    unsigned char *flags = nullptr;
    double *ex[3] = {nullptr, nullptr, nullptr};
    signed char *bkind = nullptr;
    double *bval = nullptr;             // [3 slots][3][M]
    long long bsstride = 0;             // 0, or 3M once per-substep boundary values are set
    std::vector<signed char> hbkind;    // host copy of bkind
    double *part = nullptr;   // [3][nblocks]: dt, bflux, mass
    double *chunks = nullptr; // [3][S][DE_RCHUNKS]: the reduction's first stage
    DeControl *ctl = nullptr; // [S] device step control
    unsigned long long *dneg = nullptr;
    DeControl *hctl = nullptr;          // [S] pinned host mirror of ctl
    unsigned long long *hneg = nullptr; // pinned host mirror of *dneg
    DeBatchSetup *dsetup = nullptr;     // [S] time control of a step run
    DeBatchSetup *hsetup = nullptr;     // [S] pinned staging of dsetup
    double *hist = nullptr;             // device timestep history of a step run (S == 1)
    int64_t hist_cap = 0;
    bool done_on_device = false;        // a step run may have left done flags set
    hipGraph_t graph = nullptr;         // one captured device-controlled step
    hipGraphExec_t graph_exec = nullptr;
    bool use_graph = true;

    int substate = 0;         // 0: at a step boundary (state in A); 1: rk2 substep state in B
    // per-operation calls (the reference's call order): the semi-implicit
    // momentum updates (allocated on first use), whether B holds a backup of
    // A, and the flux calls made so far (their count modulo the method's
    // substeps is the substep, as in _compute_fluxes_central)
    double *si[2] = {nullptr, nullptr};
    bool backup_valid = false;
    int64_t op_flux_calls = 0;
    // local flux updating (max_flux_update_frequency lts_M > 1): per-half-edge
    // state on the device; the reference's static counters on the host (the
    // flux call number, the step count modulo M from -1, the carried CFL
    // timestep) and its allow_timestep_increase flag
    int lts_M = 1;
    int lts_call = 0;
    long long lts_cyc = -1;
    double lts_dt = 0.0;
    int lts_allow = 1;
    DeLts lts{};
    unsigned long long neg_seen = 0;
    std::string err;

    bool timing = false;
    KernelTimer timers[T_COUNT];

    struct AdjRun *adj = nullptr;   // checkpoints of the last anuga_adjoint_record
};

// Device buffers of a recorded fixed-timestep run and of its adjoint.
struct AdjRun {
    int64_t nsteps = 0, every = 0, obs_every = 0, nckpt = 0, nobs = 0;
    int G = 0;
    double dt = 0.0;
    DeQ *ckpt = nullptr;       // [nckpt][N]: state at steps 0, every, 2 every, ...
    DeQ *seg = nullptr;        // [every][N]: the states of one segment, recomputed
    DeQ *S1 = nullptr;         // first-stage state of the step being reversed
    double *E0 = nullptr, *E1 = nullptr, *ex0 = nullptr, *ex1 = nullptr;
    unsigned char *f0 = nullptr, *f1 = nullptr;
    double *lam = nullptr, *lamn = nullptr, *lam1 = nullptr, *lamE = nullptr, *lam_ex = nullptr;
    double *lamX = nullptr;    // [16][N] the extrapolation adjoint per stencil slot
    double *zbar = nullptr, *nbar = nullptr, *rec = nullptr;
    int *gauges = nullptr;
    unsigned long long *neg = nullptr;   // fix_negative count of recomputed steps (ignored)
    ~AdjRun()
    {
        void *p[] = {ckpt, seg, S1, E0, E1, ex0, ex1, f0, f1, lam, lamn, lam1, lamE, lam_ex, lamX,
                     zbar, nbar, rec, gauges, neg};
        for (void *q : p)
            if (q) hipFree(q);
    }
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
        return best;
    };

    for (int seed = 0; seed < N; seed++) {
        if (placed[seed]) continue;
        int ecc = 0;
        int cand = bfs_last_level(seed, ecc);
        for (int it = 0; it < 8; it++) {
            int ecc2 = 0;
            int c2 = bfs_last_level(cand, ecc2);
            if (ecc2 <= ecc) break;
            cand = c2;
            ecc = ecc2;
        }
        const int root = cand;
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

// Recursive coordinate bisection of ids[lo, hi) into runs of DE_BLOCK
// triangles. Each cut is across x or y and leaves a whole number of runs on
// its first side, between 40% and 60% of them (and within 64 runs of the
// middle); of those cuts (both axes) the
// one that separates the fewest adjacent pairs of the part is taken (then the
// one nearest the middle, then x), so that the runs are compact patches whose
// boundaries follow the mesh's lines. Ties in a coordinate go by triangle
// index; each finished run is in increasing triangle index. nbr: (N,3)
// neighbours (negative: boundary), or null for cuts at the middle run. Parts
// of more than 4096 runs are cut at the middle run across their longer side
// (the search's passes over such a part cost more than they find).
// Parts are cut on several threads at once (disjoint ranges of ids and tmp).
// side[k] = 4 * (lo / DE_BLOCK + 1) + s marks triangle k as in the part
// starting at lo (a whole number of runs), s = 1 on the cut's first side and
// 2 on its second; a part reads only the marks that carry its own lo, so the
// marks of parts cut at the same time never mix.
struct TileCut {
    const double *cen;
    const int64_t *nbr;
    std::vector<int> tmp;
    std::unique_ptr<std::atomic<int>[]> side;
};

static void tile_bisect(TileCut &tc, std::vector<int> &ids, size_t lo, size_t hi, int depth)
{
    const size_t n = hi - lo;
    if (n <= (size_t)DE_BLOCK) {
        std::sort(ids.begin() + lo, ids.begin() + hi);
        return;
    }
    const double *cen = tc.cen;
    const size_t runs = (n + DE_BLOCK - 1) / DE_BLOCK;
    auto less = [cen](int dim) {
        return [cen, dim](int a, int b) {
            const double ca = cen[2 * a + dim], cb = cen[2 * b + dim];
            return ca < cb || (ca == cb && a < b);
        };
    };
    int best_dim = 0;
    size_t best_runs = runs / 2;
    if (tc.nbr && runs <= 4096) {
        // candidate cuts: 40% to 60% of the runs, at most 64 runs either side
        // of the middle (a large part finds a cut along the mesh's lines
        // within that; the window's sort stays small)
        const size_t mid = runs / 2;
        const size_t r0 = std::max<size_t>(
            {(size_t)1, (size_t)std::floor(0.4 * (double)runs), mid > 64 ? mid - 64 : (size_t)0});
        const size_t r1 = std::min<size_t>(
            {runs - 1, (size_t)std::ceil(0.6 * (double)runs), mid + 64});
        long long best_cut = -1;
        double best_off = 0.0;
        for (int dim = 0; dim < 2 && r0 <= r1; dim++) {
            int *a = tc.tmp.data() + lo;
            std::copy(ids.begin() + lo, ids.begin() + hi, a);
            const size_t pl = r0 * DE_BLOCK, ph = std::min(n, r1 * (size_t)DE_BLOCK);
            std::nth_element(a, a + pl, a + n, less(dim));
            std::nth_element(a + pl, a + ph, a + n, less(dim));
            std::sort(a + pl, a + ph, less(dim));
            const int tag = 4 * ((int)(lo / DE_BLOCK) + 1);
            auto mark = [&tc, tag](int64_t v) {
                const int m = tc.side[v].load(std::memory_order_relaxed);
                return (m & ~3) == tag ? (m & 3) : 0;
            };
            for (size_t i = 0; i < n; i++)
                tc.side[a[i]].store(tag + (i < pl ? 1 : 2), std::memory_order_relaxed);
            long long cut = 0;
            for (size_t i = 0; i < pl; i++)
                for (int j = 0; j < 3; j++) {
                    const int64_t v = tc.nbr[3 * (size_t)a[i] + j];
                    if (v >= 0 && mark(v) == 2) cut++;
                }
            for (size_t i = pl; i <= ph; i++) {
                if ((i % DE_BLOCK) == 0) {
                    const size_t r = i / DE_BLOCK;
                    const double off = std::fabs((double)r - 0.5 * (double)runs);
                    if (best_cut < 0 || cut < best_cut || (cut == best_cut && off < best_off)) {
                        best_cut = cut;
                        best_off = off;
                        best_dim = dim;
                        best_runs = r;
                    }
                }
                if (i == ph) break;
                const int e = a[i];   // moves to the first side
                for (int j = 0; j < 3; j++) {
                    const int64_t v = tc.nbr[3 * (size_t)e + j];
                    if (v < 0 || v == e) continue;
                    const int sv = mark(v);
                    if (sv == 1) cut--;
                    else if (sv == 2) cut++;
                }
                tc.side[e].store(tag + 1, std::memory_order_relaxed);
            }
            for (size_t i = 0; i < n; i++) tc.side[a[i]].store(0, std::memory_order_relaxed);
        }
    } else {
        double x0 = HUGE_VAL, y0 = HUGE_VAL, x1 = -HUGE_VAL, y1 = -HUGE_VAL;
        for (size_t i = lo; i < hi; i++) {
            const int k = ids[i];
            x0 = std::min(x0, cen[2 * k]);
            x1 = std::max(x1, cen[2 * k]);
            y0 = std::min(y0, cen[2 * k + 1]);
            y1 = std::max(y1, cen[2 * k + 1]);
        }
        best_dim = (x1 - x0 >= y1 - y0) ? 0 : 1;
    }
    const size_t nl = best_runs * (size_t)DE_BLOCK;
    std::nth_element(ids.begin() + lo, ids.begin() + lo + nl, ids.begin() + hi, less(best_dim));
    if (depth < 4 && n > ((size_t)1 << 16)) {
        // the first part on its own thread (16 at most); on one thread when
        // none can be started
        std::thread first;
        try {
            first = std::thread(tile_bisect, std::ref(tc), std::ref(ids), lo, lo + nl, depth + 1);
        } catch (...) {
            tile_bisect(tc, ids, lo, lo + nl, depth + 1);
        }
        tile_bisect(tc, ids, lo + nl, hi, depth + 1);
        if (first.joinable()) first.join();
    } else {
        tile_bisect(tc, ids, lo, lo + nl, depth + 1);
        tile_bisect(tc, ids, lo + nl, hi, depth + 1);
    }
}

// Triangles in tile order: consecutive runs of DE_BLOCK triangles are compact
// patches (the bisection above), so the halo the fused kernel extrapolates
// twice is short, and consecutive runs are neighbouring patches.
static std::vector<int> tile_order(int N, const double *cen, const int64_t *nbr)
{
    std::vector<int> o(N);
    for (int k = 0; k < N; k++) o[k] = k;
    TileCut tc{cen, nbr, std::vector<int>(nbr ? N : 0), nullptr};
    if (nbr) {
        tc.side.reset(new std::atomic<int>[N]);
        for (int k = 0; k < N; k++) tc.side[k].store(0, std::memory_order_relaxed);
    }
    tile_bisect(tc, o, 0, (size_t)N, 0);
    return o;
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
    q.low_froude = (int)p.low_froude;
    q.sloped = p.use_sloped_mannings ? 1 : 0;
    return q;
}

static int check_params(anuga_domain *d, const anuga_params *p)
{
    if (!p) return fail(d, ANUGA_E_ARG, "params is NULL");
    if (p->low_froude < 0 || p->low_froude > 2)
        return fail(d, ANUGA_E_ARG, "low_froude must be 0, 1 or 2");
    if (p->use_sloped_mannings != 0 && p->use_sloped_mannings != 1)
        return fail(d, ANUGA_E_ARG, "use_sloped_mannings must be 0 or 1");
    if (p->precision == ANUGA_PRECISION_MIXED && (p->low_froude || p->use_sloped_mannings))
        return fail(d, ANUGA_E_ARG, "low_froude and sloped friction need the FP64 path");
    if (d && p->use_sloped_mannings && !d->vx)
        return fail(d, ANUGA_E_ARG, "sloped friction needs the mesh's vertex_coordinates");
    if (p->extrapolate_velocity_second_order != 1)
        return fail(d, ANUGA_E_ARG, "extrapolate_velocity_second_order must be 1 (DE algorithms)");
    if (p->timestep_fluxcalls < 1 || p->timestep_fluxcalls > 3)
        return fail(d, ANUGA_E_ARG, "timestep_fluxcalls must be 1 (euler), 2 (rk2) or 3 (rk3)");
    if (p->precision != ANUGA_PRECISION_FP64 && p->precision != ANUGA_PRECISION_MIXED)
        return fail(d, ANUGA_E_ARG, "unknown precision");
    if (p->reorder < 0 || p->reorder > 2) return fail(d, ANUGA_E_ARG, "reorder must be 0, 1 or 2");
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
    m.scen = d->scen;
    m.nb = d->nb;
    m.cen = d->cen;
    m.gx = d->gx;
    m.gf = d->gf;
    m.area = d->area;
    m.nx = d->nx;
    m.ny = d->ny;
    m.len = d->len;
    m.radius = d->radius;
    m.rw = d->rw;
    m.rwd = d->rwd;
    m.caller = d->caller;
    m.vx = d->vx;
    return m;
}

static dim3 scenario_grid(const anuga_domain *d)
{
    return dim3((d->S + 63) / 64);
}

static void launch_extrapolate(anuga_domain *d, DeQ *S)
{
    d->e_stale = false;
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
}

static void launch_materialize(anuga_domain *d, DeQ *S)
{
    hipLaunchKernelGGL(de_materialize, dim3(d->nblocks), dim3(DE_BLOCK), 0, d->stream, d->N, S,
                       d->flags);
}

template <class T, int SUB>
static void launch_flux_t(anuga_domain *d, DeQ *S, DeQ *bk, DeQ *out)
{
    DeBoundary bc{d->bkind, d->bval, d->M, d->bsstride};
    if constexpr (sizeof(T) == 8) {
        if (d->rw || d->P.low_froude || d->P.use_sloped_mannings)
            hipLaunchKernelGGL((de_flux<double, SUB, true>), dim3(d->nblocks), dim3(DE_BLOCK), 0,
                               d->stream, dev_mesh(d), dev_params(d->P), S, d->eta,
                               (const double *)d->E, bc, d->flags, d->ex[0], d->ex[1],
                               d->ex[2], bk, out, d->ctl, d->part, d->part + d->nblocks,
                               d->dneg);
        else
            hipLaunchKernelGGL((de_flux<double, SUB>), dim3(d->nblocks), dim3(DE_BLOCK), 0,
                               d->stream, dev_mesh(d), dev_params(d->P), S, d->eta,
                               (const double *)d->E, bc, d->flags, d->ex[0], d->ex[1],
                               d->ex[2], bk, out, d->ctl, d->part, d->part + d->nblocks,
                               d->dneg);
    } else
        hipLaunchKernelGGL((de_flux_mixed<SUB>), dim3(d->nblocks), dim3(DE_BLOCK), 0, d->stream,
                           dev_mesh(d), dev_params(d->P), S, d->eta, (const DeEdgeRec32 *)d->E,
                           bc, d->flags, d->ex[0], d->ex[1], d->ex[2], bk, out, d->ctl, d->part,
                           d->part + d->nblocks, d->dneg);
}

template <class T>
static void launch_flux_p(anuga_domain *d, int substep, DeQ *S, DeQ *bk, DeQ *out)
{
    switch (substep) {
    case 0: launch_flux_t<T, 0>(d, S, bk, out); break;
    case 1: launch_flux_t<T, 1>(d, S, bk, out); break;
    case 2: launch_flux_t<T, 2>(d, S, bk, out); break;
    default: launch_flux_t<T, 3>(d, S, bk, out); break;
    }
}

// The fused extrapolation + flux kernels apply with tiles (both precisions,
// any number of scenarios, every option).
static bool fused(const anuga_domain *d)
{
    return d->hcap > 0;
}

template <int SUB>
static void launch_xflux_t(anuga_domain *d, DeQ *S, DeQ *bk, DeQ *out)
{
    DeBoundary bc{d->bkind, d->bval, d->M, d->bsstride};
    DeTiles tl{d->tstart, d->thalo, d->tslot, d->hcap, d->testart, d->tesplit, d->tedge};
    if (d->mixed) {
        // FP64 centroid depth + 12 FP32 record components per triangle
        const size_t lds = ((size_t)DE_BLOCK + (size_t)d->hcap) * (sizeof(double) + 12 * sizeof(float));
        hipLaunchKernelGGL(de_xflux_mixed<SUB>, dim3(d->nblocks), dim3(DE_BLOCK), lds, d->stream,
                           dev_mesh(d), dev_params(d->P), S, d->eta, tl, bc, d->flags, d->ex[0],
                           d->ex[1], d->ex[2], bk, out, d->ctl, d->part, d->part + d->nblocks,
                           d->part + 2 * (size_t)d->nblocks, d->dneg);
    } else if (d->rw || d->P.low_froude || d->P.use_sloped_mannings) {
        // the general instantiation: walls, low_froude, sloped friction
        const size_t lds = (12 * (size_t)DE_BLOCK + 12 * (size_t)d->hcap) * sizeof(double) +
                           (((size_t)DE_BLOCK + d->hcap + 7) & ~(size_t)7);
        hipLaunchKernelGGL((de_xflux<SUB, true>), dim3(d->nblocks), dim3(DE_BLOCK), lds,
                           d->stream, dev_mesh(d), dev_params(d->P), S, d->eta, tl, bc, d->flags,
                           d->ex[0], d->ex[1], d->ex[2], bk, out, d->ctl, d->part,
                           d->part + d->nblocks, d->part + 2 * (size_t)d->nblocks, d->dneg,
                           (double *)d->E);
    } else if (d->mirror && d->tedge && d->edge_once) {
        // each edge's flux once: tile and halo edge values, then the tile's
        // centroid depths and beds; geometry from the vertices where it can be
        const size_t lds = (14 * (size_t)DE_BLOCK + 12 * (size_t)d->hcap) * sizeof(double);
        auto kern = (d->vgeom && d->vx) ? de_xflux_e<SUB, true> : de_xflux_e<SUB, false>;
        hipLaunchKernelGGL(kern, dim3(d->nblocks), dim3(DE_BLOCK), lds, d->stream,
                           dev_mesh(d), dev_params(d->P), S, d->eta, tl, bc, d->flags, d->ex[0],
                           d->ex[1], d->ex[2], bk, out, d->ctl, d->part, d->part + d->nblocks,
                           d->part + 2 * (size_t)d->nblocks, d->dneg);
    } else {
        const size_t lds = (12 * (size_t)DE_BLOCK + 12 * (size_t)d->hcap) * sizeof(double);
        hipLaunchKernelGGL(de_xflux<SUB>, dim3(d->nblocks), dim3(DE_BLOCK), lds, d->stream,
                           dev_mesh(d), dev_params(d->P), S, d->eta, tl, bc, d->flags, d->ex[0],
                           d->ex[1], d->ex[2], bk, out, d->ctl, d->part, d->part + d->nblocks,
                           d->part + 2 * (size_t)d->nblocks, d->dneg, (double *)d->E);
    }
}

// extrapolation + flux of substep 0 (explicit updates) or 1-3 (advance S and
// combine with bk into out, de_rk_combine; out must not be S), fused over tiles
static void launch_xflux(anuga_domain *d, int substep, DeQ *S, DeQ *bk = nullptr,
                         DeQ *out = nullptr)
{
    const int T = substep ? T_XFLUX1 : T_XFLUX0;
    timer_begin(d, T);
    switch (substep) {
    case 0: launch_xflux_t<0>(d, S, bk, out); break;
    case 1: launch_xflux_t<1>(d, S, bk, out); break;
    case 2: launch_xflux_t<2>(d, S, bk, out); break;
    default: launch_xflux_t<3>(d, S, bk, out); break;
    }
    timer_end(d, T);
    d->e_stale = true;
}

// substep 0: explicit updates of S; substep > 0: advance S with the scenario's
// dt and combine with the step's initial state bk into out (de_rk_combine)
static void launch_flux(anuga_domain *d, int substep, DeQ *S, DeQ *bk = nullptr,
                        DeQ *out = nullptr)
{
    if (out == d->B) d->backup_valid = false;
    timer_begin(d, substep ? T_FLUX1 : T_FLUX0);
    if (d->mixed) launch_flux_p<float>(d, substep, S, bk, out);
    else launch_flux_p<double>(d, substep, S, bk, out);
    timer_end(d, substep ? T_FLUX1 : T_FLUX0);
}

static void launch_update(anuga_domain *d, DeQ *S, DeQ *out)
{
    if (out == d->B) d->backup_valid = false;
    timer_begin(d, T_UPDATE);
    if (d->P.use_sloped_mannings && !d->mixed)
        hipLaunchKernelGGL((de_update<true>), dim3(d->nblocks), dim3(DE_BLOCK), 0,
                           d->stream, d->scen, dev_params(d->P), S, d->eta, d->flags, d->ex[0],
                           d->ex[1], d->ex[2], d->ctl, out, d->dneg, (const double *)d->E, d->N,
                           d->vx);
    else
        hipLaunchKernelGGL(de_update<false>, dim3(d->nblocks), dim3(DE_BLOCK), 0, d->stream,
                           d->scen, dev_params(d->P), S, d->eta, d->flags, d->ex[0], d->ex[1],
                           d->ex[2], d->ctl, out, d->dneg);
    timer_end(d, T_UPDATE);
}

static void launch_reduce(anuga_domain *d, int mode)
{
    const double *pdt = d->part, *pbf = d->part + d->nblocks;
    const double *pm = d->part + 2 * (size_t)d->nblocks;
    typedef void (*Chunks)(const double *, const double *, const double *, const DeControl *,
                           double *);
    typedef void (*Final)(const double *, DeControl *);
    const Chunks kc[5] = {de_reduce_chunks<0>, de_reduce_chunks<1>, de_reduce_chunks<2>,
                          de_reduce_chunks<3>, de_reduce_chunks<4>};
    const Final kf[5] = {de_reduce_final<0>, de_reduce_final<1>, de_reduce_final<2>,
                         de_reduce_final<3>, de_reduce_final<4>};
    hipLaunchKernelGGL(kc[mode], dim3(DE_RCHUNKS, d->S), dim3(256), 0, d->stream, pdt, pbf, pm,
                       d->ctl, d->chunks);
    hipLaunchKernelGGL(kf[mode], dim3(d->S), dim3(64), 0, d->stream, d->chunks, d->ctl);
}

static void launch_control(anuga_domain *d, void (*k)(DeControl *, int))
{
    hipLaunchKernelGGL(k, scenario_grid(d), dim3(64), 0, d->stream, d->ctl, d->S);
}

static void launch_volume(anuga_domain *d, DeQ *S)
{
    hipLaunchKernelGGL(de_volume, dim3(d->nblocks), dim3(DE_BLOCK), 0, d->stream, d->scen, d->ctl,
                       S, d->area, d->part + 2 * (size_t)d->nblocks);
}

// leave device-controlled mode before host-controlled launches
static void ensure_live(anuga_domain *d)
{
    if (d->done_on_device) {
        launch_control(d, de_clear_done);
        d->done_on_device = false;
    }
}

// One device-controlled step of every scenario: both substeps, the timestep
// on the device and the end-of-step bookkeeping. Every kernel returns at once
// for a scenario whose done flag is set.
static void enqueue_later_substeps(anuga_domain *d)
{
    if (d->P.timestep_fluxcalls == 2) {
        if (fused(d)) {
            launch_xflux(d, 1, d->B, d->A, d->A);
        } else {
            launch_extrapolate(d, d->B);
            launch_flux(d, 1, d->B, d->A, d->A);
        }
        launch_reduce(d, 1);
    } else if (d->P.timestep_fluxcalls == 3) {
        if (fused(d) && d->C) {
            // the middle substep writes C, not B in place: other tiles are
            // still extrapolating B
            launch_xflux(d, 2, d->B, d->A, d->C);
            launch_reduce(d, 1);
            launch_xflux(d, 3, d->C, d->A, d->A);
            launch_reduce(d, 4);
            return;
        }
        launch_extrapolate(d, d->B);
        launch_flux(d, 2, d->B, d->A, d->B);
        launch_reduce(d, 1);
        launch_extrapolate(d, d->B);
        launch_flux(d, 3, d->B, d->A, d->A);
        launch_reduce(d, 4);
    }
}

// fractional-step operators after a step (apply_fractional_steps,
// generic_domain.py:1849, :2312-2314), in the order they were added
static void launch_operator(anuga_domain *d, const RateOp &o)
{
    if (o.n == 0) return;   // an empty region
    const dim3 g((o.n + DE_BLOCK - 1) / DE_BLOCK), b(DE_BLOCK);
    if (o.kind == 2) {
        DeInlet in{o.n, o.idx, o.rate, o.total_area, o.Q1, o.Q2, o.vmode, o.u, o.v, o.influx};
        hipLaunchKernelGGL(de_inlet, dim3(1), dim3(DE_INLET_THREADS), 0, d->stream, d->ctl, in,
                           d->A);
    } else if (o.kind == 1)
        hipLaunchKernelGGL(de_set_stage, g, b, 0, d->stream, d->ctl, o.n, o.idx, o.rate, o.rate0,
                           d->A);
    else
        hipLaunchKernelGGL(de_rate, g, b, 0, d->stream, d->ctl, o.n, o.idx, o.rate, o.rate0,
                           o.factor, o.careful, d->A, d->area, o.influx);
}

static void launch_operators(anuga_domain *d)
{
    for (const RateOp &o : d->ops) launch_operator(d, o);
}

static void enqueue_device_step(anuga_domain *d)
{
    const bool rk = d->P.timestep_fluxcalls > 1;
    if (fused(d)) {
        launch_xflux(d, 0, d->A);
    } else {
        launch_extrapolate(d, d->A);
        launch_flux(d, 0, d->A);
    }
    launch_reduce(d, 0);
    launch_control(d, de_timestep);
    launch_update(d, d->A, rk ? d->B : d->A);
    enqueue_later_substeps(d);
    launch_operators(d);
    hipLaunchKernelGGL(de_finish_step, scenario_grid(d), dim3(64), 0, d->stream, d->ctl, d->S,
                       d->S == 1 ? d->hist : nullptr);
    if (d->G)
        hipLaunchKernelGGL(de_sample_gauges<false>, dim3(1), dim3(256), 0, d->stream, d->ctl,
                           d->G, d->gtri, d->A, 0.0, d->gtimes, d->gvals);
}

static void drop_graph(anuga_domain *d)
{
    if (d->graph_exec) hipGraphExecDestroy(d->graph_exec);
    if (d->graph) hipGraphDestroy(d->graph);
    d->graph_exec = nullptr;
    d->graph = nullptr;
}

// read the control blocks and counters; raise a pending device error
static int sync_scalars(anuga_domain *d)
{
    HIPCHK(d, hipMemcpyAsync(d->hctl, d->ctl, d->S * sizeof(DeControl), hipMemcpyDeviceToHost,
                             d->stream));
    HIPCHK(d, hipMemcpyAsync(d->hneg, d->dneg, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             d->stream));
    HIPCHK(d, hipStreamSynchronize(d->stream));
    int e = 0;
    for (int s = 0; s < d->S; s++) e |= d->hctl[s].status;
    if (e) {
        launch_control(d, de_clear_status);
        hipStreamSynchronize(d->stream);
        for (int s = 0; s < d->S; s++) d->hctl[s].status = 0;
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
// creation
// ---------------------------------------------------------------------------
extern "C" void anuga_destroy(anuga_domain *d);

// Allocate the buffers of S scenarios of ntri[s] triangles and nbnd[s]
// boundary edges; no mesh is uploaded yet. On failure *out stays NULL and
// the message is the create error.
static int domain_alloc(const anuga_params *params, int64_t S64, const int64_t *ntri,
                        const int64_t *nbnd, anuga_domain **out)
{
    *out = nullptr;
    int rc = check_params(nullptr, params);
    if (rc) return rc;
    if (S64 < 1 || S64 >= (int64_t(1) << 24)) return fail(nullptr, ANUGA_E_ARG, "bad scenario count");
    const int S = (int)S64;
    int64_t total = 0, mtotal = 0;
    std::vector<int> kbeg(S), mbeg(S);
    for (int s = 0; s < S; s++) {
        if (ntri[s] <= 0 || ntri[s] >= (int64_t(1) << 29))
            return fail(nullptr, ANUGA_E_ARG, "number_of_elements must be in [1, 2^29)");
        if (nbnd[s] < 0 || nbnd[s] >= (int64_t(1) << 30))
            return fail(nullptr, ANUGA_E_ARG, "bad boundary_length");
        // every scenario starts on a workgroup boundary
        total = (total + DE_BLOCK - 1) / DE_BLOCK * DE_BLOCK;
        kbeg[s] = (int)total;
        mbeg[s] = (int)mtotal;
        total += ntri[s];
        mtotal += nbnd[s];
    }
    if (total >= (int64_t(1) << 29) || mtotal >= (int64_t(1) << 30))
        return fail(nullptr, ANUGA_E_ARG, "too many triangles for one handle (2^29)");

    anuga_domain *d = new anuga_domain();
    d->S = S;
    d->N = (int)total;
    d->M = (int)mtotal;
    d->P = *params;
    d->nblocks = (d->N + DE_BLOCK - 1) / DE_BLOCK;
    d->kbeg = kbeg;
    d->mbeg = mbeg;
    d->ntri.resize(S);
    d->nbnd.resize(S);
    for (int s = 0; s < S; s++) {
        d->ntri[s] = (int)ntri[s];
        d->nbnd[s] = (int)nbnd[s];
    }
    d->mesh_set.assign(S, 0);
    d->meshes_missing = S;
    // a scenario without boundary edges needs no boundary conditions
    d->bnd_set.assign(S, 0);
    d->boundaries_missing = 0;
    for (int s = 0; s < S; s++) {
        d->bnd_set[s] = d->nbnd[s] == 0;
        if (!d->bnd_set[s]) d->boundaries_missing++;
    }
    d->perm.assign(d->N, -1);
    const int N = d->N, M = d->M;

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
    DCHK(dalloc(&d->bval, 9 * (size_t)M));   // three substep slots
    DCHK(dalloc(&d->part, 3 * (size_t)d->nblocks));
    DCHK(dalloc(&d->chunks, 3 * (size_t)S * DE_RCHUNKS));
    DCHK(dalloc(&d->ctl, S));
    DCHK(dalloc(&d->dsetup, S));
    DCHK(dalloc(&d->dneg, 1));
    DCHK(hipHostMalloc((void **)&d->hctl, S * sizeof(DeControl), hipHostMallocDefault));
    DCHK(hipHostMalloc((void **)&d->hsetup, S * sizeof(DeBatchSetup), hipHostMallocDefault));
    DCHK(hipHostMalloc((void **)&d->hneg, sizeof(unsigned long long), hipHostMallocDefault));
    if (S > 1) {
        DCHK(dalloc(&d->scen, d->nblocks));
        std::vector<int> bs(d->nblocks, 0);
        for (int s = 0; s < S; s++) {
            const int b0 = d->kbeg[s] / DE_BLOCK, b1 = (d->kbeg[s] + d->ntri[s] + DE_BLOCK - 1) / DE_BLOCK;
            for (int b = b0; b < b1; b++) bs[b] = s;
        }
        DCHK(hipMemcpy(d->scen, bs.data(), d->nblocks * sizeof(int), hipMemcpyHostToDevice));
    }

    // zero state, geometry and partials; the padding triangles between
    // scenarios stay zero and no kernel computes on them
    DCHK(hipMemset(d->nb, 0, N * sizeof(int4)));
    DCHK(hipMemset(d->cen, 0, N * sizeof(double2)));
    DCHK(hipMemset(d->area, 0, N * sizeof(double)));
    DCHK(hipMemset(d->gx, 0, N * (d->mixed ? sizeof(DeGeoX32) : sizeof(DeGeoX64))));
    if (d->mixed) {
        DCHK(hipMemset(d->gf, 0, N * sizeof(DeGeoF32)));
    } else {
        DCHK(hipMemset(d->nx, 0, n3 * sizeof(double)));
        DCHK(hipMemset(d->ny, 0, n3 * sizeof(double)));
        DCHK(hipMemset(d->len, 0, n3 * sizeof(double)));
        DCHK(hipMemset(d->radius, 0, N * sizeof(double)));
    }
    DCHK(hipMemset(d->eta, 0, N * sizeof(double)));
    DCHK(hipMemset(d->A, 0, N * sizeof(DeQ)));
    DCHK(hipMemset(d->B, 0, N * sizeof(DeQ)));
    for (int q = 0; q < 3; q++) DCHK(hipMemset(d->ex[q], 0, N * sizeof(double)));
    DCHK(hipMemset(d->E, 0, d->mixed ? N * sizeof(DeEdgeRec32) : 4 * n3 * sizeof(double)));
    DCHK(hipMemset(d->flags, DE_F_KEEP, N));
    DCHK(hipMemset(d->bkind, 0, std::max(M, 1)));
    DCHK(hipMemset(d->bval, 0, 9 * (size_t)std::max(M, 1) * sizeof(double)));
    DCHK(hipMemset(d->part, 0, 3 * (size_t)d->nblocks * sizeof(double)));
    DCHK(hipMemset(d->dneg, 0, sizeof(unsigned long long)));
    std::memset(d->hctl, 0, S * sizeof(DeControl));
    std::memset(d->hsetup, 0, S * sizeof(DeBatchSetup));
    for (int s = 0; s < S; s++) {
        DeControl &c = d->hctl[s];
        c.kbeg = d->kbeg[s];
        c.kend = d->kbeg[s] + d->ntri[s];
        c.bbeg = d->kbeg[s] / DE_BLOCK;
        c.bend = (c.kend + DE_BLOCK - 1) / DE_BLOCK;
    }
    DCHK(hipMemcpy(d->ctl, d->hctl, S * sizeof(DeControl), hipMemcpyHostToDevice));
    *d->hneg = 0;
    {
        const char *g = std::getenv("ANUGA_HIP_GRAPH");
        d->use_graph = !(g && g[0] == '0');
        const char *e = std::getenv("ANUGA_HIP_EDGE_ONCE");
        d->edge_once = !(e && e[0] == '0');
        const char *vg = std::getenv("ANUGA_HIP_VERTEX_GEOMETRY");
        d->vgeom = !(vg && vg[0] == '0') && !d->mixed;
    }
    DCHK(hipDeviceSynchronize());
#undef DCHK
    *out = d;
    return ANUGA_OK;
}

// Check a mesh of N triangles and M boundary edges (the reference's Mesh
// arrays) on the host. Failures are reported on d (the create error if NULL).
static int validate_mesh(anuga_domain *d, const anuga_mesh *mesh, int64_t N, int64_t M)
{
    if (!mesh) return fail(d, ANUGA_E_ARG, "mesh is NULL");
    if (mesh->number_of_elements != N || mesh->boundary_length != M)
        return fail(d, ANUGA_E_ARG, "mesh size differs from the size given at creation");
    const int64_t *nbr = mesh->neighbours, *nbe = mesh->neighbour_edges;
    const int64_t *sur = mesh->surrogate_neighbours, *nob = mesh->number_of_boundaries;
    if (!nbr || !nbe || !sur || !nob || !mesh->normals || !mesh->edgelengths || !mesh->radii ||
        !mesh->areas || !mesh->centroid_coordinates || !mesh->edge_coordinates)
        return fail(d, ANUGA_E_ARG, "a mesh array is NULL");

    // connectivity: the device derives surrogate neighbours and the boundary
    // count from `neighbours`, which requires no internal boundaries
    for (int64_t k = 0; k < N; k++) {
        int nb = 0;
        for (int i = 0; i < 3; i++) {
            const int64_t ki = 3 * k + i, n = nbr[ki];
            if (n >= 0) {
                if (n >= N) return fail(d, ANUGA_E_ARG, "neighbour index out of range");
                const int64_t m = nbe[ki];
                if (m < 0 || m > 2 || nbr[3 * n + m] != k)
                    return fail(d, ANUGA_E_ARG, "inconsistent neighbour_edges");
                if (sur[ki] != n)
                    return fail(d, ANUGA_E_ARG, "internal boundary edges are not supported");
            } else {
                if (-n - 1 >= M) return fail(d, ANUGA_E_ARG, "boundary index out of range");
                if (sur[ki] != k)
                    return fail(d, ANUGA_E_ARG, "internal boundary edges are not supported");
                nb++;
            }
        }
        if (nob[k] != nb) return fail(d, ANUGA_E_ARG, "number_of_boundaries inconsistent");
    }
    return ANUGA_OK;
}

// Halo lists of the tiles of one scenario (internal numbering, nb records hn):
// for tile b (triangles [b DE_BLOCK, (b + 1) DE_BLOCK)), the neighbour
// triangles outside the tile, each once, in order of first appearance; per
// triangle, each edge's slot in its tile's list (1023: the neighbour is in
// the tile). No tiles (hcap 0) if a halo exceeds kMaxHalo
// so that every fused kernel's LDS stays within 64 KB per workgroup: 96 B
// per halo entry and up to 113 B per tile triangle (the once-per-edge
// kernel's edge values, centroid depths and beds; the general
// instantiation's flag byte), the static reduction arrays, and the halo
// stride rounded up to a multiple of 32.
static const int kMaxHalo = std::min(320, (65536 - 113 * DE_BLOCK - 1024) / 97 / 32 * 32);

// The tile metadata of triangles in internal order (nb: packed neighbours of
// one scenario starting at 0). Returns the largest halo, or -1 when a
// neighbour index is out of range; the caller refuses halos above kMaxHalo.
// The tile edge lists of the once-per-edge fused kernel (de_xflux_e): per
// tile, each edge between two tile triangles from its lower-numbered side and
// each edge to the halo from the tile's side, ordered by edge index, then
// triangle (consecutive entries read consecutive geometry). halo slots as
// tile_halos assigns them.
static void tile_edges(int N, const int4 *hn, const std::vector<unsigned> &slot,
                       std::vector<int> &estart, std::vector<int> &esplit,
                       std::vector<unsigned> &edges)
{
    const int nt = (N + DE_BLOCK - 1) / DE_BLOCK;
    estart.assign(nt + 1, 0);
    esplit.assign(nt, 0);
    edges.clear();
    edges.reserve((size_t)N * 5 / 3 + 64);
    for (int b = 0; b < nt; b++) {
        estart[b] = (int)edges.size();
        const int k0 = b * DE_BLOCK, k1 = std::min(N, k0 + DE_BLOCK);
        // the edges inside the tile first, then the edges to the halo
        for (int halo = 0; halo < 2; halo++) {
            if (halo) esplit[b] = (int)edges.size();
            for (int i = 0; i < 3; i++)
                for (int k = k0; k < k1; k++) {
                    const int v = i == 0 ? hn[k].x : (i == 1 ? hn[k].y : hn[k].z);
                    if (v < 0) continue;
                    const int n = v >> 2, j = v & 3;
                    const unsigned s = (slot[k] >> (10 * i)) & 1023u;
                    if (halo && s != 1023u)
                        edges.push_back(DE_TE_MAKE(k - k0, i, 1, s, j));
                    else if (!halo && s == 1023u && k < n)
                        edges.push_back(DE_TE_MAKE(k - k0, i, 0, n - k0, j));
                }
        }
    }
    estart[nt] = (int)edges.size();
}

static int tile_halos(int N, const int4 *hn, std::vector<int> &start, std::vector<int> &halo,
                      std::vector<unsigned> &slot)
{
    const int nt = (N + DE_BLOCK - 1) / DE_BLOCK;
    start.assign(nt + 1, 0);
    halo.clear();
    halo.reserve((size_t)nt * 64);
    slot.assign(N, 0);
    std::vector<int> seen_tile(N, -1), seen_slot(N, 0);
    int hcap = 0;
    for (int b = 0; b < nt; b++) {
        start[b] = (int)halo.size();
        const int k0 = b * DE_BLOCK, k1 = std::min(N, k0 + DE_BLOCK);
        for (int k = k0; k < k1; k++) {
            const int v[3] = {hn[k].x, hn[k].y, hn[k].z};
            unsigned sl = 0;
            for (int i = 0; i < 3; i++) {
                unsigned si = 1023u;
                if (v[i] >= 0) {
                    const int n = v[i] >> 2;
                    if (n >= N || (v[i] & 3) == 3) return -1;
                    if (n < k0 || n >= k1) {
                        if (seen_tile[n] != b) {
                            seen_tile[n] = b;
                            seen_slot[n] = (int)(halo.size() - start[b]);
                            halo.push_back(n);
                        }
                        si = (unsigned)std::min(seen_slot[n], 1023);
                    }
                }
                sl |= si << (10 * i);
            }
            slot[k] = sl;
        }
        hcap = std::max(hcap, (int)halo.size() - start[b]);
    }
    start[nt] = (int)halo.size();
    return hcap;
}

// The tile metadata of scenario s (its nb records hn, internal ids), kept on
// the host until every scenario's mesh is set.
static int scenario_tiles(anuga_domain *d, int s, const std::vector<int4> &hn)
{
    const int N = (int)hn.size(), kb = d->kbeg[s];
    std::vector<int4> loc(hn);
    for (int4 &r : loc) {
        int *v[3] = {&r.x, &r.y, &r.z};
        for (int i = 0; i < 3; i++)
            if (*v[i] >= 0) *v[i] -= kb << 2;
    }
    if (d->ts_hcap.empty()) {
        d->ts_start.resize(d->S);
        d->ts_halo.resize(d->S);
        d->ts_slot.resize(d->S);
        d->ts_estart.resize(d->S);
        d->ts_esplit.resize(d->S);
        d->ts_edge.resize(d->S);
        d->ts_hcap.assign(d->S, -1);
    }
    const int hcap = tile_halos(N, loc.data(), d->ts_start[s], d->ts_halo[s], d->ts_slot[s]);
    if (hcap < 0) return fail(d, ANUGA_E_ARG, "tiles: neighbour out of range");
    tile_edges(N, loc.data(), d->ts_slot[s], d->ts_estart[s], d->ts_esplit[s], d->ts_edge[s]);
    for (int &n : d->ts_halo[s]) n += kb;
    d->ts_hcap[s] = hcap;
    return ANUGA_OK;
}

// Assemble every scenario's tile metadata into the device arrays the fused
// kernels read (block b's halo at tstart[b]); no tiles (hcap 0) if a scenario
// has none or a halo exceeds kMaxHalo.
static int upload_tiles(anuga_domain *d)
{
    d->hcap = 0;
    for (int **p : {&d->tstart, &d->thalo}) {
        if (*p) HIPCHK(d, hipFree(*p));
        *p = nullptr;
    }
    if (d->tslot) HIPCHK(d, hipFree(d->tslot));
    d->tslot = nullptr;
    if (d->testart) HIPCHK(d, hipFree(d->testart));
    if (d->tesplit) HIPCHK(d, hipFree(d->tesplit));
    if (d->tedge) HIPCHK(d, hipFree(d->tedge));
    d->testart = d->tesplit = nullptr;
    d->tedge = nullptr;
    if ((int)d->ts_hcap.size() != d->S) return ANUGA_OK;
    int hcap = 0;
    for (int s = 0; s < d->S; s++) {
        if (d->ts_hcap[s] < 0) return ANUGA_OK;
        hcap = std::max(hcap, d->ts_hcap[s]);
    }
    if (hcap > kMaxHalo) return ANUGA_OK;
    std::vector<int> start(d->nblocks + 1, 0), halo, estart(d->nblocks + 1, 0),
        esplit(d->nblocks, 0);
    std::vector<unsigned> slot(d->N, 0x3fffffffu), edge;
    int b = 0;
    for (int s = 0; s < d->S; s++) {
        const int b0 = d->kbeg[s] / DE_BLOCK, nt = (int)d->ts_start[s].size() - 1;
        for (; b < b0; b++) {
            start[b] = (int)halo.size();
            estart[b] = esplit[b] = (int)edge.size();
        }
        for (int t = 0; t < nt; t++, b++) {
            start[b] = (int)halo.size() + d->ts_start[s][t];
            estart[b] = (int)edge.size() + d->ts_estart[s][t];
            esplit[b] = (int)edge.size() + d->ts_esplit[s][t];
        }
        halo.insert(halo.end(), d->ts_halo[s].begin(), d->ts_halo[s].end());
        edge.insert(edge.end(), d->ts_edge[s].begin(), d->ts_edge[s].end());
        std::copy(d->ts_slot[s].begin(), d->ts_slot[s].end(), slot.begin() + d->kbeg[s]);
    }
    for (; b <= d->nblocks; b++) {
        start[b] = (int)halo.size();
        estart[b] = (int)edge.size();
        if (b < d->nblocks) esplit[b] = (int)edge.size();
    }
    HIPCHK(d, dalloc(&d->testart, (size_t)d->nblocks + 1));
    HIPCHK(d, dalloc(&d->tesplit, (size_t)d->nblocks));
    HIPCHK(d, hipMemcpy(d->tesplit, esplit.data(), d->nblocks * sizeof(int),
                        hipMemcpyHostToDevice));
    HIPCHK(d, dalloc(&d->tedge, std::max<size_t>(edge.size(), 1)));
    HIPCHK(d, hipMemcpy(d->testart, estart.data(), (d->nblocks + 1) * sizeof(int),
                        hipMemcpyHostToDevice));
    if (!edge.empty())
        HIPCHK(d, hipMemcpy(d->tedge, edge.data(), edge.size() * sizeof(unsigned),
                            hipMemcpyHostToDevice));
    HIPCHK(d, dalloc(&d->tstart, (size_t)d->nblocks + 1));
    HIPCHK(d, dalloc(&d->thalo, halo.size()));
    HIPCHK(d, dalloc(&d->tslot, (size_t)d->N));
    HIPCHK(d, hipMemcpy(d->tstart, start.data(), (d->nblocks + 1) * sizeof(int),
                        hipMemcpyHostToDevice));
    if (!halo.empty())
        HIPCHK(d, hipMemcpy(d->thalo, halo.data(), halo.size() * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(d, hipMemcpy(d->tslot, slot.data(), d->N * sizeof(unsigned), hipMemcpyHostToDevice));
    // the LDS stride: a multiple of 32, never zero
    d->hcap = std::max((hcap + 31) / 32 * 32, 32);
    return ANUGA_OK;
}

// Upload the validated mesh of scenario s (its own triangle and boundary
// numbering, as the reference's Mesh holds it).
// Whether each interior edge's normal is exactly the negation of its
// neighbour's and the lengths are equal (true of normals and lengths formed
// from the two vertices, general_mesh.py:202-232), so that one side's flux
// gives the other's by negation.
static bool mesh_mirrors(const anuga_mesh *mesh)
{
    const int64_t N = mesh->number_of_elements;
    const double *nrm = mesh->normals, *el = mesh->edgelengths;
    for (int64_t k = 0; k < N; k++)
        for (int i = 0; i < 3; i++) {
            const int64_t n = mesh->neighbours[3 * k + i];
            if (n < 0) continue;
            const int64_t j = mesh->neighbour_edges[3 * k + i];
            if (!(nrm[6 * n + 2 * j] == -nrm[6 * k + 2 * i]) ||
                !(nrm[6 * n + 2 * j + 1] == -nrm[6 * k + 2 * i + 1]) ||
                !(el[3 * n + j] == el[3 * k + i]))
                return false;
        }
    return true;
}

// Whether the mesh's edge midpoints, normals and lengths are the ones
// general_mesh.py forms from its vertex coordinates (:202-232, :599-626),
// bit for bit, so that the kernels may form them from the vertices.
static bool geometry_from_vertices(const anuga_mesh *mesh)
{
    const double *V = mesh->vertex_coordinates;
    if (!V) return false;
    const double *nrm = mesh->normals, *el = mesh->edgelengths, *em = mesh->edge_coordinates;
    for (int64_t k = 0; k < mesh->number_of_elements; k++) {
        const double *v = V + 6 * k;
        for (int i = 0; i < 3; i++) {
            const int a = (i + 1) % 3, b = (i + 2) % 3;
            const double mx = 0.5 * (v[2 * a] + v[2 * b]), my = 0.5 * (v[2 * a + 1] + v[2 * b + 1]);
            if (!(mx == em[6 * k + 2 * i]) || !(my == em[6 * k + 2 * i + 1])) return false;
            const double xn = v[2 * b] - v[2 * a], yn = v[2 * b + 1] - v[2 * a + 1];
            const double l = std::sqrt(xn * xn + yn * yn);
            if (!(l == el[3 * k + i]) || !(yn / l == nrm[6 * k + 2 * i]) ||
                !(-(xn / l) == nrm[6 * k + 2 * i + 1]))
                return false;
        }
    }
    return true;
}

// The vertex geometry stays on only while every scenario's arrays are
// formed from its vertices (the kernels then read the per-triangle vertex
// coordinates, vx).
static void check_vertex_geometry(anuga_domain *d, const anuga_mesh *mesh)
{
    if (!geometry_from_vertices(mesh)) d->vgeom = false;
}

static int upload_mesh(anuga_domain *d, int s, const anuga_mesh *mesh)
{
    if (!mesh_mirrors(mesh)) d->mirror = false;
    const int N = d->ntri[s], kb = d->kbeg[s], mb = d->mbeg[s];
    const int64_t *nbr = mesh->neighbours, *nbe = mesh->neighbour_edges;
    const int64_t *sur = mesh->surrogate_neighbours, *nob = mesh->number_of_boundaries;
    std::vector<int> perm(N), iperm(N);
    if (d->P.reorder == 2) perm = tile_order(N, mesh->centroid_coordinates, mesh->neighbours);
    else if (d->P.reorder) perm = rcm_order(N, nbr);
    else
        for (int k = 0; k < N; k++) perm[k] = k;
    for (int p = 0; p < N; p++) iperm[perm[p]] = p;

    const double *nrm = mesh->normals, *el = mesh->edgelengths;
    const double *cen = mesh->centroid_coordinates, *em = mesh->edge_coordinates;
    std::vector<int4> hn(N);
    std::vector<double2> hc(N);
    std::vector<double> ha(N);
    for (int p = 0; p < N; p++) {
        const int64_t k = perm[p];
        int v[3];
        for (int i = 0; i < 3; i++) {
            const int64_t n = nbr[3 * k + i];
            v[i] = n >= 0 ? (((kb + iperm[n]) << 2) | (int)nbe[3 * k + i]) : (int)(n - mb);
        }
        hn[p] = make_int4(v[0], v[1], v[2], (int)nob[k]);
        hc[p] = make_double2(cen[2 * k], cen[2 * k + 1]);
        ha[p] = mesh->areas[k];
    }
    HIPCHK(d, hipStreamSynchronize(d->stream));
    HIPCHK(d, hipMemcpy(d->nb + kb, hn.data(), N * sizeof(int4), hipMemcpyHostToDevice));
    if (d->hnb.empty()) d->hnb.assign(d->N, make_int4(-1, -1, -1, 0));
    std::copy(hn.begin(), hn.end(), d->hnb.begin() + kb);
    HIPCHK(d, hipMemcpy(d->cen + kb, hc.data(), N * sizeof(double2), hipMemcpyHostToDevice));
    HIPCHK(d, hipMemcpy(d->area + kb, ha.data(), N * sizeof(double), hipMemcpyHostToDevice));
    if (!d->mixed) {
        const size_t NN = d->N;
        std::vector<DeGeoX64> gx(N);
        std::vector<double> a(3 * (size_t)N), b(3 * (size_t)N), c(3 * (size_t)N), r(N);
        for (int p = 0; p < N; p++) {
            const int64_t k = perm[p];
            for (int i = 0; i < 3; i++) {
                // xv - x and yv - y exactly as sw_domain_orig.c:2130-2135 forms them
                gx[p].dxv[i] = em[6 * k + 2 * i] - cen[2 * k];
                gx[p].dyv[i] = em[6 * k + 2 * i + 1] - cen[2 * k + 1];
                a[(size_t)i * N + p] = nrm[6 * k + 2 * i];
                b[(size_t)i * N + p] = nrm[6 * k + 2 * i + 1];
                c[(size_t)i * N + p] = el[3 * k + i];
            }
            r[p] = mesh->radii[k];
            // the gradient's auxiliary geometry, in the kernel's operation order
            // (this file is compiled without floating-point contraction)
            double *g = gx[p].aux;
            for (int j = 0; j < 6; j++) g[j] = 0.0;
            if (nob[k] <= 1) {
                int64_t s[3];
                for (int i = 0; i < 3; i++) s[i] = sur[3 * k + i];
                const double dx1 = cen[2 * s[1]] - cen[2 * s[0]];
                const double dx2 = cen[2 * s[2]] - cen[2 * s[0]];
                const double dy1 = cen[2 * s[1] + 1] - cen[2 * s[0] + 1];
                const double dy2 = cen[2 * s[2] + 1] - cen[2 * s[0] + 1];
                const double area2 = dy2 * dx1 - dy1 * dx2;
                g[0] = dx1;
                g[1] = dx2;
                g[2] = dy1;
                g[3] = dy2;
                g[4] = area2;
                g[5] = 1.0 / area2;
            } else if (nob[k] == 2) {
                int64_t s1 = -1;
                for (int i = 0; i < 3 && s1 < 0; i++)
                    if (nbr[3 * k + i] >= 0) s1 = nbr[3 * k + i];
                const double dx1 = cen[2 * s1] - cen[2 * k];
                const double dy1 = cen[2 * s1 + 1] - cen[2 * k + 1];
                const double area2 = dx1 * dx1 + dy1 * dy1;
                double dx2 = 1.0 / area2;
                const double dy2 = dx2 * dy1;
                dx2 *= dx1;
                g[1] = dx2;
                g[3] = dy2;
            }
        }
        HIPCHK(d, hipMemcpy((DeGeoX64 *)d->gx + kb, gx.data(), N * sizeof(DeGeoX64),
                            hipMemcpyHostToDevice));
        for (int i = 0; i < 3; i++) {
            const size_t o = (size_t)i * NN + kb, h = (size_t)i * N;
            HIPCHK(d, hipMemcpy(d->nx + o, a.data() + h, N * sizeof(double), hipMemcpyHostToDevice));
            HIPCHK(d, hipMemcpy(d->ny + o, b.data() + h, N * sizeof(double), hipMemcpyHostToDevice));
            HIPCHK(d, hipMemcpy(d->len + o, c.data() + h, N * sizeof(double), hipMemcpyHostToDevice));
        }
        HIPCHK(d, hipMemcpy(d->radius + kb, r.data(), N * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<DeGeoX32> gx(N);
        std::vector<DeGeoF32> gf(N);
        for (int p = 0; p < N; p++) {
            const int64_t k = perm[p];
            for (int i = 0; i < 3; i++) {
                gx[p].dxv[i] = (float)(em[6 * k + 2 * i] - cen[2 * k]);
                gx[p].dyv[i] = (float)(em[6 * k + 2 * i + 1] - cen[2 * k + 1]);
                const int64_t sn = sur[3 * k + i];
                gx[p].rx[i] = (float)(cen[2 * sn] - cen[2 * k]);
                gx[p].ry[i] = (float)(cen[2 * sn + 1] - cen[2 * k + 1]);
                gf[p].nx[i] = (float)nrm[6 * k + 2 * i];
                gf[p].ny[i] = (float)nrm[6 * k + 2 * i + 1];
                gf[p].len[i] = (float)el[3 * k + i];
                gf[p].dzc[i] = gf[p].pad[i] = 0.f;
            }
            gf[p].radius = (float)mesh->radii[k];
        }
        HIPCHK(d, hipMemcpy((DeGeoX32 *)d->gx + kb, gx.data(), N * sizeof(DeGeoX32),
                            hipMemcpyHostToDevice));
        HIPCHK(d, hipMemcpy((DeGeoF32 *)d->gf + kb, gf.data(), N * sizeof(DeGeoF32),
                            hipMemcpyHostToDevice));
    }
    std::copy(perm.begin(), perm.end(), d->perm.begin() + kb);
    if (!d->caller) HIPCHK(d, dalloc(&d->caller, (size_t)d->N));
    HIPCHK(d, hipMemcpy(d->caller + kb, perm.data(), N * sizeof(int), hipMemcpyHostToDevice));
    if (mesh->vertex_coordinates && !d->mixed) {
        std::vector<double> hv(6 * (size_t)N);
        for (int p = 0; p < N; p++)
            for (int j = 0; j < 6; j++) hv[6 * (size_t)p + j] = mesh->vertex_coordinates[6 * (size_t)perm[p] + j];
        if (!d->vx) HIPCHK(d, dalloc(&d->vx, 6 * (size_t)d->N));
        HIPCHK(d, hipMemcpy(d->vx + 6 * (size_t)kb, hv.data(), 6 * (size_t)N * sizeof(double),
                            hipMemcpyHostToDevice));
    }
    if (d->vgeom) check_vertex_geometry(d, mesh);
    if (d->P.reorder == 2) {
        int rc = scenario_tiles(d, s, hn);
        if (rc) return rc;
    }
    if (!d->mesh_set[s]) {
        d->mesh_set[s] = 1;
        d->meshes_missing--;
    }
    if (d->P.reorder == 2 && d->meshes_missing == 0) return upload_tiles(d);
    return ANUGA_OK;
}

// ---------------------------------------------------------------------------
// per-scenario state access
// ---------------------------------------------------------------------------
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

static int download_records(anuga_domain *d, const DeQ *src, int kb, int n, std::vector<DeQ> &h)
{
    h.resize(n);
    HIPCHK(d, hipStreamSynchronize(d->stream));
    HIPCHK(d, hipMemcpy(h.data(), src + kb, n * sizeof(DeQ), hipMemcpyDeviceToHost));
    return ANUGA_OK;
}

static int scenario_ok(anuga_domain *d, int64_t s)
{
    if (s < 0 || s >= d->S) return fail(d, ANUGA_E_ARG, "scenario index out of range");
    if (!d->mesh_set[s]) return fail(d, ANUGA_E_STATE, "the scenario's mesh is not set");
    return ANUGA_OK;
}

static int set_centroids(anuga_domain *d, int s, int quantity, const double *values)
{
    if (!values) return fail(d, ANUGA_E_ARG, "values is NULL");
    const int N = d->ntri[s], kb = d->kbeg[s];
    const int *perm = d->perm.data() + kb;
    if (quantity == ANUGA_FRICTION) {
        std::vector<double> t(N);
        for (int p = 0; p < N; p++) t[p] = values[perm[p]];
        HIPCHK(d, hipStreamSynchronize(d->stream));
        HIPCHK(d, hipMemcpy(d->eta + kb, t.data(), N * sizeof(double), hipMemcpyHostToDevice));
        return ANUGA_OK;
    }
    const int slot = record_slot(quantity);
    if (slot < 0) return fail(d, ANUGA_E_ARG, "quantity cannot be set");
    // the bed is part of both state buffers; the others of the current one
    DeQ *bufs[2] = {d->substate ? d->B : d->A, slot == 3 ? (d->substate ? d->A : d->B) : nullptr};
    for (DeQ *buf : bufs) {
        if (!buf) continue;
        std::vector<DeQ> h;
        int rc = download_records(d, buf, kb, N, h);
        if (rc) return rc;
        for (int p = 0; p < N; p++) (&h[p].w)[slot] = values[perm[p]];
        HIPCHK(d, hipMemcpy(buf + kb, h.data(), N * sizeof(DeQ), hipMemcpyHostToDevice));
    }
    if (slot == 3 && d->mixed) {
        // bed offsets of the mixed flux's local datum
        hipLaunchKernelGGL(de_bed_offsets, dim3((N + DE_BLOCK - 1) / DE_BLOCK), dim3(DE_BLOCK), 0,
                           d->stream, kb, kb + N, d->nb, d->A, (DeGeoF32 *)d->gf);
        int rc = launch_check(d);
        if (rc) return rc;
        HIPCHK(d, hipStreamSynchronize(d->stream));
    }
    return ANUGA_OK;
}

static int get_centroids(anuga_domain *d, int s, int quantity, double *values)
{
    if (!values) return fail(d, ANUGA_E_ARG, "values is NULL");
    const int N = d->ntri[s], kb = d->kbeg[s];
    const int *perm = d->perm.data() + kb;
    if (quantity == ANUGA_FRICTION) {
        std::vector<double> t(N);
        HIPCHK(d, hipStreamSynchronize(d->stream));
        HIPCHK(d, hipMemcpy(t.data(), d->eta + kb, N * sizeof(double), hipMemcpyDeviceToHost));
        for (int p = 0; p < N; p++) values[perm[p]] = t[p];
        return ANUGA_OK;
    }
    const int slot = record_slot(quantity);
    if (slot < 0 && quantity != ANUGA_HEIGHT) return fail(d, ANUGA_E_ARG, "unknown quantity");
    std::vector<DeQ> h;
    int rc = download_records(d, d->substate ? d->B : d->A, kb, N, h);
    if (rc) return rc;
    for (int p = 0; p < N; p++)
        values[perm[p]] = slot >= 0 ? (&h[p].w)[slot] : std::fmax(h[p].w - h[p].z, 0.0);
    return ANUGA_OK;
}

static int set_boundary(anuga_domain *d, int s, const int32_t *kind, const double *dirichlet_values)
{
    const int M = d->nbnd[s], mb = d->mbeg[s], MT = d->M;
    if (M > 0 && !kind) return fail(d, ANUGA_E_ARG, "kind is NULL");
    if (M == 0) return ANUGA_OK;
    std::vector<signed char> k(M, 0);
    std::vector<double> v(3 * (size_t)M, 0.0);
    for (int j = 0; j < M; j++) {
        if (kind[j] < 0 || kind[j] > 6) return fail(d, ANUGA_E_ARG, "unknown boundary kind");
        k[j] = (signed char)kind[j];
        if (kind[j] == ANUGA_BC_DIRICHLET || kind[j] == ANUGA_BC_SET_STAGE_NORMAL_MOMENTUM ||
            kind[j] == ANUGA_BC_SET_STAGE_TRANSMISSIVE_MOMENTUM) {
            if (!dirichlet_values) return fail(d, ANUGA_E_ARG, "boundary values are NULL");
            for (int q = 0; q < 3; q++) v[(size_t)q * M + j] = dirichlet_values[3 * (size_t)j + q];
        }
    }
    HIPCHK(d, hipStreamSynchronize(d->stream));
    HIPCHK(d, hipMemcpy(d->bkind + mb, k.data(), M, hipMemcpyHostToDevice));
    if (d->hbkind.empty()) d->hbkind.assign(std::max(MT, 1), 0);
    std::copy(k.begin(), k.end(), d->hbkind.begin() + mb);
    for (int slot = 0; slot < 3; slot++)
        for (int q = 0; q < 3; q++)
            HIPCHK(d, hipMemcpy(d->bval + ((size_t)slot * 3 + q) * MT + mb,
                                v.data() + (size_t)q * M, M * sizeof(double),
                                hipMemcpyHostToDevice));
    if (d->bsstride) drop_graph(d);
    d->bsstride = 0;
    if (!d->bnd_set[s]) {
        d->bnd_set[s] = 1;
        d->boundaries_missing--;
    }
    return ANUGA_OK;
}

static int ready_to_step(anuga_domain *d)
{
    if (d->meshes_missing) return fail(d, ANUGA_E_STATE, "a scenario's mesh is not set");
    if (d->boundaries_missing) return fail(d, ANUGA_E_STATE, "boundary conditions not set");
    if (d->P.use_sloped_mannings && !d->vx)
        return fail(d, ANUGA_E_ARG, "sloped friction needs the mesh's vertex_coordinates");
    // rk3 with the fused kernels: a third state for the middle substep
    // (allocated here, before any graph capture)
    if (d->P.timestep_fluxcalls == 3 && fused(d) && !d->C) {
        HIPCHK(d, hipStreamSynchronize(d->stream));
        HIPCHK(d, dalloc(&d->C, (size_t)d->N));
    }
    return ANUGA_OK;
}

static int single(anuga_domain *d)
{
    if (d->S != 1)
        return fail(d, ANUGA_E_STATE, "an ensemble handle: use the anuga_ensemble_* entry points");
    return ANUGA_OK;
}

// Up to max_steps device-controlled steps of every scenario with time control
// tc[s] and state st[s] (see anuga_evolve_steps).
static int evolve_steps(anuga_domain *d, int64_t max_steps, const anuga_time_control *tc,
                        anuga_step_state *st, double *dt_history, bool stop_reached)
{
    if (d->substate != 0) return fail(d, ANUGA_E_STATE, "evolve_steps called inside an rk2 step");
    int rc = ready_to_step(d);
    if (rc) return rc;
    if (d->bsstride)
        return fail(d, ANUGA_E_STATE, "boundary values set per substep time need host-sequenced "
                                      "steps (anuga_compute_fluxes / anuga_update_...)");
    if (d->lts_M > 1)
        return fail(d, ANUGA_E_STATE, "local flux updating needs host-sequenced steps "
                                      "(anuga_compute_flux_update_frequency after each timestep)");
    if (d->G) {   // every step records a gauge row: stop where the rows run out
        if (d->grows_host >= d->gcap)
            return fail(d, ANUGA_E_STATE, "gauge rows full: download them first");
        max_steps = std::min<int64_t>(max_steps, d->gcap - d->grows_host);
    }
    if (dt_history && max_steps > d->hist_cap) {
        if (d->hist) hipFree(d->hist);
        d->hist = nullptr;
        d->hist_cap = 0;
        drop_graph(d);
        HIPCHK(d, dalloc(&d->hist, (size_t)max_steps));
        d->hist_cap = max_steps;
    }
    HIPCHK(d, hipStreamSynchronize(d->stream));   // the staging buffer is free again
    for (int s = 0; s < d->S; s++) {
        DeBatchSetup &b = d->hsetup[s];
        b.t = st[s].relative_time;
        b.rec_min = st[s].recorded_min_timestep;
        b.rec_max = st[s].recorded_max_timestep;
        b.cfl = tc[s].CFL;
        b.max_dt = tc[s].evolve_max_timestep;
        b.min_dt = tc[s].evolve_min_timestep;
        b.finaltime = tc[s].relative_finaltime;
        b.yieldtime = tc[s].relative_yieldtime;
        b.fixed_flux_dt = tc[s].fixed_flux_timestep;
        b.eps = tc[s].epsilon;
        b.smallsteps = (int)st[s].smallsteps;
        b.order = (int)st[s].order;
        b.default_order = (int)tc[s].default_order;
        b.max_smallsteps = (int)tc[s].max_smallsteps;
        b.hist_cap = dt_history ? (int)std::min<int64_t>(max_steps, d->hist_cap) : 0;
        b.method = (int)d->P.timestep_fluxcalls;
        b.stop_reached = stop_reached ? 1 : 0;
    }
    HIPCHK(d, hipMemcpyAsync(d->dsetup, d->hsetup, d->S * sizeof(DeBatchSetup),
                             hipMemcpyHostToDevice, d->stream));
    hipLaunchKernelGGL(de_setup_batch, scenario_grid(d), dim3(64), 0, d->stream, d->ctl, d->S,
                       d->dsetup);
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
    rc = launch_check(d);
    if (rc) return rc;
    rc = sync_scalars(d);
    for (int s = 0; s < d->S; s++) {
        const DeControl &c = d->hctl[s];
        st[s].relative_time = c.t;
        st[s].timestep = c.dt;
        st[s].flux_timestep = c.flux_dt;
        st[s].recorded_min_timestep = c.rec_min;
        st[s].recorded_max_timestep = c.rec_max;
        st[s].smallsteps = c.smallsteps;
        st[s].order = c.order;
        st[s].steps_taken = c.nsteps;
        st[s].reached = c.done;
    }
    if (d->G) d->grows_host += d->hctl[0].nsteps;
    if (dt_history && d->hctl[0].nsteps > 0) {
        HIPCHK(d, hipMemcpy(dt_history, d->hist,
                            (size_t)std::min<int64_t>(d->hctl[0].nsteps, max_steps) * sizeof(double),
                            hipMemcpyDeviceToHost));
    }
    return rc;
}

// protect + extrapolation of every scenario, written back to the centroids
static int distribute(anuga_domain *d)
{
    if (d->meshes_missing) return fail(d, ANUGA_E_STATE, "a scenario's mesh is not set");
    ensure_live(d);
    DeQ *S = d->substate ? d->B : d->A;
    launch_extrapolate(d, S);
    launch_materialize(d, S);
    launch_reduce(d, 2);
    int rc = launch_check(d);
    if (rc) return rc;
    return sync_scalars(d);
}

static int volumes(anuga_domain *d)
{
    DeQ *S = d->substate ? d->B : d->A;
    launch_volume(d, S);
    launch_reduce(d, 3);
    int rc = launch_check(d);
    if (rc) return rc;
    return sync_scalars(d);
}

// ---------------------------------------------------------------------------
// fixed-timestep runs and their adjoint
// ---------------------------------------------------------------------------
// limiter jumps below this size (m, m/s) count as rounding noise in the
// adjoint: the limiter's factor is held there (de_adjoint.h adj_limit)
const double kAdjLimiterFreeze = 1.0e-10;

static AdjView adj_view(const anuga_domain *d, double dt)
{
    AdjView m;
    m.N = d->N;
    m.M = d->M;
    m.nb = reinterpret_cast<const int *>(d->nb);
    m.cen = reinterpret_cast<const double *>(d->cen);
    m.gx = reinterpret_cast<const double *>(d->gx);
    m.gx_stride = sizeof(DeGeoX64) / sizeof(double);
    m.nx = d->nx;
    m.ny = d->ny;
    m.len = d->len;
    m.area = d->area;
    m.eta = d->eta;
    m.bkind = d->bkind;
    m.bval = d->bval;
    const DeParams q = dev_params(d->P);
    m.p = AdjParams{q.g, q.eps, q.mah, q.bw, q.bwd, q.buh, q.buhd, q.bvh, q.bvhd,
                    kAdjLimiterFreeze};
    m.dt = dt;
    return m;
}

// one DE1 step of X in place with the timestep already in ctl->dt (the
// production kernels; d->B, d->E, d->flags and d->ex serve as scratch)
static void step_fixed(anuga_domain *d, DeQ *X)
{
    if (fused(d)) {
        launch_xflux(d, 0, X);
        launch_update(d, X, d->B);
        launch_xflux(d, 1, d->B, X, X);
        return;
    }
    launch_extrapolate(d, X);
    launch_flux(d, 0, X);
    launch_update(d, X, d->B);
    launch_extrapolate(d, d->B);
    launch_flux(d, 1, d->B, X, X);
}

// the stage quantities of the step from X into the adjoint's buffers
static void adj_stages(anuga_domain *d, AdjRun *a, const DeQ *X)
{
    const dim3 g(d->nblocks), b(DE_BLOCK);
    const size_t N = d->N;
    DeBoundary bc{d->bkind, d->bval, d->M, 0};
    const DeParams p = dev_params(d->P);
    double *pm = d->part + 2 * (size_t)d->nblocks;
    hipLaunchKernelGGL(de_extrapolate<double>, g, b, 0, d->stream, dev_mesh(d), p, X, a->E0, a->f0,
                       pm, d->ctl);
    hipLaunchKernelGGL((de_flux<double, 0>), g, b, 0, d->stream, dev_mesh(d), p, X, d->eta,
                       a->E0, bc, a->f0, a->ex0, a->ex0 + N, a->ex0 + 2 * N, nullptr, nullptr,
                       d->ctl,
                       d->part, d->part + d->nblocks, a->neg);
    hipLaunchKernelGGL(de_update<false>, g, b, 0, d->stream, d->scen, p, X, d->eta, a->f0, a->ex0,
                       a->ex0 + N, a->ex0 + 2 * N, d->ctl, a->S1, a->neg);
    hipLaunchKernelGGL(de_extrapolate<double>, g, b, 0, d->stream, dev_mesh(d), p, a->S1, a->E1,
                       a->f1, pm, d->ctl);
    hipLaunchKernelGGL((de_flux<double, 0>), g, b, 0, d->stream, dev_mesh(d), p, a->S1, d->eta,
                       a->E1, bc, a->f1, a->ex1, a->ex1 + N, a->ex1 + 2 * N, nullptr, nullptr,
                       d->ctl,
                       d->part, d->part + d->nblocks, a->neg);
}

// lam (adjoint of the step's result, [3][N]) -> adjoint of its start X
static int adj_step(anuga_domain *d, AdjRun *a, const DeQ *X)
{
    const dim3 g(d->nblocks), b(DE_BLOCK);
    const size_t N = d->N;
    const AdjView m = adj_view(d, a->dt);
    const double *Sx = reinterpret_cast<const double *>(X);
    const double *S1 = reinterpret_cast<const double *>(a->S1);
    adj_stages(d, a, X);
    // the final saxpy: half to the step's start directly, half through the stages
    hipLaunchKernelGGL(adj_k_scale, dim3((3 * N + DE_BLOCK - 1) / DE_BLOCK), b, 0, d->stream,
                       (int)(3 * N), 0.5, a->lam, a->lamn);
    HIPCHK(d, hipMemsetAsync(a->lam1, 0, 3 * N * sizeof(double), d->stream));
    // second stage: S2 = advance(S1)
    hipLaunchKernelGGL(adj_k_advance, g, b, 0, d->stream, m, S1, a->f1, a->ex1, a->ex1 + N,
                       a->ex1 + 2 * N, a->lam, 0.5, a->lam1, a->zbar, a->nbar, a->lam_ex);
    HIPCHK(d, hipMemsetAsync(a->lamE, 0, 12 * N * sizeof(double), d->stream));
    hipLaunchKernelGGL(adj_k_flux, g, b, 0, d->stream, m, S1, a->f1, a->E1, a->lam_ex, a->lamE,
                       a->lam1, a->zbar);
    hipLaunchKernelGGL(adj_k_extrap, g, b, 0, d->stream, m, S1, a->f1, a->lamE, a->lamX);
    hipLaunchKernelGGL(adj_k_extrap_gather, g, b, 0, d->stream, m, a->lamX, a->lam1, a->zbar);
    // first stage: S1 = advance(X)
    hipLaunchKernelGGL(adj_k_advance, g, b, 0, d->stream, m, Sx, a->f0, a->ex0, a->ex0 + N,
                       a->ex0 + 2 * N, a->lam1, 1.0, a->lamn, a->zbar, a->nbar, a->lam_ex);
    HIPCHK(d, hipMemsetAsync(a->lamE, 0, 12 * N * sizeof(double), d->stream));
    hipLaunchKernelGGL(adj_k_flux, g, b, 0, d->stream, m, Sx, a->f0, a->E0, a->lam_ex, a->lamE,
                       a->lamn, a->zbar);
    hipLaunchKernelGGL(adj_k_extrap, g, b, 0, d->stream, m, Sx, a->f0, a->lamE, a->lamX);
    hipLaunchKernelGGL(adj_k_extrap_gather, g, b, 0, d->stream, m, a->lamX, a->lamn, a->zbar);
    std::swap(a->lam, a->lamn);
    return launch_check(d);
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
    const int64_t n = mesh->number_of_elements, m = mesh->boundary_length;
    int rc = check_params(nullptr, params);
    if (rc) return rc;
    if (n <= 0 || n >= (int64_t(1) << 29))
        return fail(nullptr, ANUGA_E_ARG, "number_of_elements must be in [1, 2^29)");
    if (m < 0 || m >= (int64_t(1) << 30)) return fail(nullptr, ANUGA_E_ARG, "bad boundary_length");
    rc = validate_mesh(nullptr, mesh, n, m);
    if (rc) return rc;
    anuga_domain *d = nullptr;
    rc = domain_alloc(params, 1, &n, &m, &d);
    if (rc) return rc;
    rc = upload_mesh(d, 0, mesh);
    if (rc) {
        const std::string msg = d->err;
        anuga_destroy(d);
        return fail(nullptr, rc, msg);
    }
    *out = d;
    return ANUGA_OK;
}

void anuga_destroy(anuga_domain *d)
{
    if (!d) return;
    if (d->stream) hipStreamSynchronize(d->stream);
    drop_graph(d);
    void *ptrs[] = {d->nb, d->cen, d->gx, d->gf, d->area, d->nx, d->ny, d->len, d->radius,
                    d->scen, d->tstart, d->thalo, d->tslot, d->rw, d->rwd, d->caller, d->vx, d->node_start, d->node_tri, d->iperm_dev,
                    d->out_f32, d->gtri, d->gtimes, d->gvals, d->sub_idx, d->sub_val, d->eta, d->A, d->B, d->C, d->E, d->flags,
                    d->ex[0], d->ex[1], d->ex[2], d->bkind, d->bval, d->part, d->chunks, d->ctl, d->dsetup,
                    d->dneg, d->hist, d->si[0], d->si[1], d->testart, d->tesplit, d->tedge,
                    d->lts.e0, d->lts.e1, d->lts.e2, d->lts.pg, d->lts.et, d->lts.fuf,
                    d->lts.acf, d->lts.tmin, d->lts.unf};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (d->hctl) hipHostFree(d->hctl);
    if (d->hsetup) hipHostFree(d->hsetup);
    if (d->hneg) hipHostFree(d->hneg);
    for (int t = 0; t < T_COUNT; t++) {
        for (auto e : d->timers[t].start) hipEventDestroy(e);
        for (auto e : d->timers[t].stop) hipEventDestroy(e);
    }
    delete d->adj;
    for (RateOp &o : d->ops) {
        void *q[] = {o.idx, o.rate, o.influx};
        for (void *x : q)
            if (x) hipFree(x);
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
    if (d->lts_M > 1 && params->timestep_fluxcalls != 1)
        return fail(d, ANUGA_E_STATE,
                    "Local extrapolation and flux updating only supported with euler timestepping");
    if (d->lts_M > 1 && (params->low_froude || params->use_sloped_mannings))
        return fail(d, ANUGA_E_STATE, "local flux updating runs without low_froude or sloped "
                                      "friction");
    if (std::memcmp(params, &d->P, sizeof(anuga_params)) != 0) drop_graph(d);
    // a new method restarts the per-operation flux substep count, as the
    // reference's static base_call does (sw_domain_orig.c:791-794)
    if (params->timestep_fluxcalls != d->P.timestep_fluxcalls) d->op_flux_calls = 0;
    d->P = *params;
    return ANUGA_OK;
}

int anuga_set_centroid_values(anuga_domain *d, int quantity, const double *values)
{
    if (!d || !values) return ANUGA_E_ARG;
    int rc = single(d);
    return rc ? rc : set_centroids(d, 0, quantity, values);
}

int anuga_get_centroid_values(anuga_domain *d, int quantity, double *values)
{
    if (!d || !values) return ANUGA_E_ARG;
    int rc = single(d);
    return rc ? rc : get_centroids(d, 0, quantity, values);
}

int anuga_get_edge_values(anuga_domain *d, int quantity, double *values)
{
    if (!d || !values) return ANUGA_E_ARG;
    int rc = single(d);
    if (rc) return rc;
    const int N = d->N;
    const size_t n3 = 3 * (size_t)N;
    if (quantity != ANUGA_STAGE && quantity != ANUGA_HEIGHT && quantity != ANUGA_XMOMENTUM &&
        quantity != ANUGA_YMOMENTUM && quantity != ANUGA_ELEVATION)
        return fail(d, ANUGA_E_ARG, "unknown edge quantity");
    if (d->e_stale) {
        // the last steps ran the fused kernel, whose edge values stay in LDS:
        // extrapolate the current state (what distribute_to_vertices_and_edges
        // computes, without writing protect's outcome back)
        if (d->meshes_missing) return fail(d, ANUGA_E_STATE, "a scenario's mesh is not set");
        launch_extrapolate(d, d->substate ? d->B : d->A);
        rc = launch_check(d);
        if (rc) return rc;
    }
    HIPCHK(d, hipStreamSynchronize(d->stream));
    std::vector<double> e(4 * n3);
    if (d->mixed) {
        // edge = centroid + jump; the centroid state is the one the last
        // extrapolation saw after it was written back (distribute)
        std::vector<DeEdgeRec32> r(N);
        std::vector<DeQ> st;
        HIPCHK(d, hipMemcpy(r.data(), d->E, N * sizeof(DeEdgeRec32), hipMemcpyDeviceToHost));
        rc = download_records(d, d->substate ? d->B : d->A, 0, N, st);
        if (rc) return rc;
        for (int p = 0; p < N; p++)
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
        for (int p = 0; p < N; p++)
            for (int i = 0; i < 3; i++)
                for (int qq = 0; qq < 4; qq++)
                    e[4 * (3 * (size_t)p + i) + qq] = soa[qq * n3 + (size_t)i * N + p];
    }
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
    int rc = single(d);
    return rc ? rc : set_boundary(d, 0, kind, dirichlet_values);
}

int anuga_set_boundary_values(anuga_domain *d, int64_t slot, const double *values)
{
    if (!d) return ANUGA_E_ARG;
    int rc = single(d);
    if (rc) return rc;
    const int M = d->M;
    if (slot < 0 || slot > 2) return fail(d, ANUGA_E_ARG, "slot must be 0, 1 or 2");
    if (M > 0 && !values) return fail(d, ANUGA_E_ARG, "values is NULL");
    if (M == 0) return ANUGA_OK;
    std::vector<double> v(3 * (size_t)M);
    for (int j = 0; j < M; j++)
        for (int q = 0; q < 3; q++) v[(size_t)q * M + j] = values[3 * (size_t)j + q];
    HIPCHK(d, hipStreamSynchronize(d->stream));
    HIPCHK(d, hipMemcpy(d->bval + (size_t)slot * 3 * M, v.data(), 3 * (size_t)M * sizeof(double),
                        hipMemcpyHostToDevice));
    if (slot > 0 && d->bsstride == 0) {
        d->bsstride = 3 * (long long)M;
        drop_graph(d);
    }
    return ANUGA_OK;
}

static int upload_rate(anuga_domain *d, RateOp &o, const double *rates, double rate,
                       double factor)
{
    o.rate0 = rate;
    o.factor = factor;
    if (rates) {
        if (!o.rate) HIPCHK(d, dalloc(&o.rate, (size_t)o.n));
        HIPCHK(d, hipMemcpy(o.rate, rates, o.n * sizeof(double), hipMemcpyHostToDevice));
        o.careful = 0;
        for (int i = 0; i < o.n; i++) o.careful |= rates[i] < 0.0;
    } else {
        if (o.rate) HIPCHK(d, hipFree(o.rate));
        o.rate = nullptr;
        o.careful = rate < 0.0;
    }
    return ANUGA_OK;
}

static int add_operator(anuga_domain *d, int kind, int64_t n, const int64_t *indices,
                        const double *rates, double rate, double factor, int64_t *id);

int anuga_add_rate_operator(anuga_domain *d, int64_t n, const int64_t *indices,
                            const double *rates, double rate, double factor, int64_t *id)
{
    return add_operator(d, 0, n, indices, rates, rate, factor, id);
}

int anuga_add_set_stage_operator(anuga_domain *d, int64_t n, const int64_t *indices,
                                 const double *stages, double stage, int64_t *id)
{
    return add_operator(d, 1, n, indices, stages, stage, 1.0, id);
}

static int add_operator(anuga_domain *d, int kind, int64_t n, const int64_t *indices,
                        const double *rates, double rate, double factor, int64_t *id)
{
    if (!d || !id) return ANUGA_E_ARG;
    int rc = single(d);
    if (rc) return rc;
    const int N = d->ntri[0];
    if (indices && (n < 0 || n > N)) return fail(d, ANUGA_E_ARG, "bad operator triangle count");
    RateOp o;
    o.kind = kind;
    o.n = indices ? (int)n : N;
    std::vector<int> ip(N), idx(std::max(o.n, 1));
    for (int p = 0; p < N; p++) ip[d->perm[p]] = p;
    for (int i = 0; i < o.n; i++) {
        const int64_t t = indices ? indices[i] : i;
        if (t < 0 || t >= N) return fail(d, ANUGA_E_ARG, "operator triangle out of range");
        idx[i] = ip[t];
    }
    HIPCHK(d, hipStreamSynchronize(d->stream));
    HIPCHK(d, dalloc(&o.idx, (size_t)std::max(o.n, 1)));
    HIPCHK(d, dalloc(&o.influx, 1));
    HIPCHK(d, hipMemcpy(o.idx, idx.data(), std::max(o.n, 1) * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(d, hipMemset(o.influx, 0, sizeof(double)));
    d->ops.push_back(o);
    rc = upload_rate(d, d->ops.back(), rates, rate, factor);
    if (rc) return rc;
    drop_graph(d);
    *id = (int64_t)d->ops.size() - 1;
    return ANUGA_OK;
}

int anuga_add_inlet_operator(anuga_domain *d, int64_t n, const int64_t *indices,
                             const double *areas, double total_area, int64_t velocity_mode,
                             double u, double v, double Q1, double Q2, int64_t *id)
{
    if (!d || !id || !indices || !areas || n < 1) return ANUGA_E_ARG;
    int rc = single(d);
    if (rc) return rc;
    if (n > DE_INLET_MAX)
        return fail(d, ANUGA_E_ARG, "inlet regions on the device hold at most 2048 triangles");
    if (velocity_mode < 0 || velocity_mode > 2) return fail(d, ANUGA_E_ARG, "bad velocity mode");
    const int N = d->ntri[0];
    RateOp o;
    o.kind = 2;
    o.n = (int)n;
    std::vector<int> ip(N), idx(n);
    for (int p = 0; p < N; p++) ip[d->perm[p]] = p;
    for (int64_t i = 0; i < n; i++) {
        if (indices[i] < 0 || indices[i] >= N)
            return fail(d, ANUGA_E_ARG, "operator triangle out of range");
        idx[i] = ip[indices[i]];
    }
    o.total_area = total_area;
    o.Q1 = Q1;
    o.Q2 = Q2;
    o.vmode = (int)velocity_mode;
    o.u = u;
    o.v = v;
    HIPCHK(d, hipStreamSynchronize(d->stream));
    HIPCHK(d, dalloc(&o.idx, (size_t)n));
    HIPCHK(d, dalloc(&o.rate, (size_t)n));
    HIPCHK(d, dalloc(&o.influx, 4));
    HIPCHK(d, hipMemcpy(o.idx, idx.data(), n * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(d, hipMemcpy(o.rate, areas, n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(d, hipMemset(o.influx, 0, 4 * sizeof(double)));
    d->ops.push_back(o);
    drop_graph(d);
    *id = (int64_t)d->ops.size() - 1;
    return ANUGA_OK;
}

int anuga_update_inlet_operator(anuga_domain *d, int64_t id, double Q1, double Q2)
{
    if (!d) return ANUGA_E_ARG;
    if (id < 0 || id >= (int64_t)d->ops.size() || d->ops[id].kind != 2)
        return fail(d, ANUGA_E_ARG, "no such inlet operator");
    if (d->ops[id].Q1 != Q1 || d->ops[id].Q2 != Q2) drop_graph(d);
    d->ops[id].Q1 = Q1;
    d->ops[id].Q2 = Q2;
    return ANUGA_OK;
}

int anuga_get_inlet_operator_stats(anuga_domain *d, int64_t id, double *stats)
{
    if (!d || !stats) return ANUGA_E_ARG;
    if (id < 0 || id >= (int64_t)d->ops.size() || d->ops[id].kind != 2)
        return fail(d, ANUGA_E_ARG, "no such inlet operator");
    HIPCHK(d, hipMemcpyAsync(stats, d->ops[id].influx, 4 * sizeof(double), hipMemcpyDeviceToHost,
                             d->stream));
    HIPCHK(d, hipStreamSynchronize(d->stream));
    return ANUGA_OK;
}

int anuga_update_rate_operator(anuga_domain *d, int64_t id, const double *rates, double rate,
                               double factor)
{
    if (!d) return ANUGA_E_ARG;
    if (id < 0 || id >= (int64_t)d->ops.size()) return fail(d, ANUGA_E_ARG, "no such operator");
    if (d->ops[id].kind == 2)
        return fail(d, ANUGA_E_ARG, "an inlet operator: use anuga_update_inlet_operator");
    HIPCHK(d, hipStreamSynchronize(d->stream));
    RateOp &o = d->ops[id];
    const bool had_array = o.rate != nullptr;
    const RateOp before = o;
    int rc = upload_rate(d, o, rates, rate, factor);
    if (rc) return rc;
    if (had_array != (o.rate != nullptr) || before.rate0 != o.rate0 ||
        before.factor != o.factor || before.careful != o.careful)
        drop_graph(d);
    return ANUGA_OK;
}

// internal ids of n caller triangles in d->sub_idx (staging grown as needed)
static int stage_subset(anuga_domain *d, int64_t n, const int64_t *indices)
{
    const int N = d->ntri[0];
    if (n < 0 || n > N) return fail(d, ANUGA_E_ARG, "bad subset size");
    if (n > 0 && !indices) return fail(d, ANUGA_E_ARG, "indices is NULL");
    if (d->hiperm.empty()) {
        d->hiperm.resize(N);
        for (int p = 0; p < N; p++) d->hiperm[d->perm[p]] = p;
    }
    std::vector<int> idx(std::max<int64_t>(n, 1));
    for (int64_t i = 0; i < n; i++) {
        if (indices[i] < 0 || indices[i] >= N) return fail(d, ANUGA_E_ARG, "index out of range");
        idx[i] = d->hiperm[indices[i]];
    }
    HIPCHK(d, hipStreamSynchronize(d->stream));
    if (n > d->sub_cap) {
        if (d->sub_idx) HIPCHK(d, hipFree(d->sub_idx));
        if (d->sub_val) HIPCHK(d, hipFree(d->sub_val));
        d->sub_idx = nullptr;
        d->sub_val = nullptr;
        HIPCHK(d, dalloc(&d->sub_idx, (size_t)n));
        HIPCHK(d, dalloc(&d->sub_val, (size_t)n));
        d->sub_cap = (int)n;
    }
    if (n > 0)
        HIPCHK(d, hipMemcpy(d->sub_idx, idx.data(), n * sizeof(int), hipMemcpyHostToDevice));
    return ANUGA_OK;
}

int anuga_get_centroid_subset(anuga_domain *d, int64_t n, const int64_t *indices, double *values)
{
    if (!d) return ANUGA_E_ARG;
    int rc = single(d);
    if (rc) return rc;
    if (n > 0 && !values) return fail(d, ANUGA_E_ARG, "values is NULL");
    if ((rc = stage_subset(d, n, indices))) return rc;
    if (n == 0) return ANUGA_OK;
    hipLaunchKernelGGL(de_gather_state, dim3((n + DE_BLOCK - 1) / DE_BLOCK), dim3(DE_BLOCK), 0,
                       d->stream, (int)n, d->sub_idx, d->substate ? d->B : d->A, d->sub_val);
    rc = launch_check(d);
    if (rc) return rc;
    HIPCHK(d, hipMemcpyAsync(values, d->sub_val, n * sizeof(DeQ), hipMemcpyDeviceToHost,
                             d->stream));
    HIPCHK(d, hipStreamSynchronize(d->stream));
    return ANUGA_OK;
}

int anuga_set_centroid_subset(anuga_domain *d, int64_t n, const int64_t *indices,
                              const double *values)
{
    if (!d) return ANUGA_E_ARG;
    int rc = single(d);
    if (rc) return rc;
    if (d->substate != 0) return fail(d, ANUGA_E_STATE, "inside a multi-stage step");
    if (n > 0 && !values) return fail(d, ANUGA_E_ARG, "values is NULL");
    if ((rc = stage_subset(d, n, indices))) return rc;
    if (n == 0) return ANUGA_OK;
    std::vector<int64_t> sorted(indices, indices + n);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
        return fail(d, ANUGA_E_ARG, "indices must be distinct");
    HIPCHK(d, hipMemcpy(d->sub_val, values, n * sizeof(DeQ), hipMemcpyHostToDevice));
    ensure_live(d);
    hipLaunchKernelGGL(de_scatter_state, dim3((n + DE_BLOCK - 1) / DE_BLOCK), dim3(DE_BLOCK), 0,
                       d->stream, (int)n, d->sub_idx, d->sub_val, d->A);
    rc = launch_check(d);
    if (rc) return rc;
    HIPCHK(d, hipStreamSynchronize(d->stream));
    return ANUGA_OK;
}

int anuga_apply_operator(anuga_domain *d, int64_t id)
{
    if (!d) return ANUGA_E_ARG;
    if (id < 0 || id >= (int64_t)d->ops.size()) return fail(d, ANUGA_E_ARG, "no such operator");
    if (d->substate != 0) return fail(d, ANUGA_E_STATE, "inside a multi-stage step");
    ensure_live(d);
    launch_operator(d, d->ops[id]);
    return launch_check(d);
}

int anuga_apply_operators(anuga_domain *d)
{
    if (!d) return ANUGA_E_ARG;
    int rc = single(d);
    if (rc) return rc;
    if (d->substate != 0) return fail(d, ANUGA_E_STATE, "inside a multi-stage step");
    ensure_live(d);
    launch_operators(d);
    return launch_check(d);
}

int anuga_get_operator_influx(anuga_domain *d, int64_t id, double *cumulative)
{
    if (!d || !cumulative) return ANUGA_E_ARG;
    if (id < 0 || id >= (int64_t)d->ops.size()) return fail(d, ANUGA_E_ARG, "no such operator");
    HIPCHK(d, hipStreamSynchronize(d->stream));
    HIPCHK(d, hipMemcpy(cumulative, d->ops[id].influx, sizeof(double), hipMemcpyDeviceToHost));
    return ANUGA_OK;
}

int anuga_set_riverwalls(anuga_domain *d, const int64_t *edge_flux_type,
                         int64_t number_of_riverwall_edges, const double *riverwall_elevation,
                         const int64_t *riverwall_rowIndex, int64_t nrows, int64_t ncols,
                         const double *hydraulic_properties)
{
    if (!d) return ANUGA_E_ARG;
    int rc = single(d);
    if (rc) return rc;
    if (d->substate != 0) return fail(d, ANUGA_E_STATE, "inside a multi-stage step");
    if (d->lts_M > 1 && number_of_riverwall_edges > 0)
        return fail(d, ANUGA_E_STATE, "local flux updating runs without riverwalls");
    const int N = d->ntri[0];
    const int64_t R = number_of_riverwall_edges;
    if (R < 0 || R > 3 * (int64_t)N) return fail(d, ANUGA_E_ARG, "bad riverwall edge count");
    if (R > 0) {
        if (d->mixed) return fail(d, ANUGA_E_ARG, "riverwalls need the FP64 path");
        if (!edge_flux_type || !riverwall_elevation || !riverwall_rowIndex ||
            !hydraulic_properties)
            return fail(d, ANUGA_E_ARG, "a riverwall array is NULL");
        if (ncols < 5 || nrows < 1)
            return fail(d, ANUGA_E_ARG, "the hydraulic table needs >= 1 row of >= 5 columns");
    }
    // ranks of the flagged half-edges in (triangle, edge) order
    std::vector<int> rank(3 * (size_t)N, -1);
    int64_t r = 0;
    if (R > 0)
        for (int64_t ki = 0; ki < 3 * (int64_t)N; ki++)
            if (edge_flux_type[ki] == 1) {
                if (r == R) return fail(d, ANUGA_E_ARG, "more flagged edges than riverwall edges");
                rank[ki] = (int)r++;
            }
    if (r != R) return fail(d, ANUGA_E_ARG, "fewer flagged edges than riverwall edges");
    std::vector<DeRiverwall> rec(R);
    for (int64_t j = 0; j < R; j++) {
        const int64_t row = riverwall_rowIndex[j];
        if (row < 0 || row >= nrows) return fail(d, ANUGA_E_ARG, "riverwall_rowIndex out of range");
        const double *h = hydraulic_properties + row * ncols;
        rec[j] = DeRiverwall{riverwall_elevation[j], h[0], h[1], h[2], h[3], h[4]};
    }
    // An interior edge is computed once, from the half-edge of the triangle
    // visited first, and mirrored (sw_domain_orig.c:822-840, :1023-1037): it
    // is a wall iff that half-edge is flagged, with that half-edge's record.
    // Both half-edges take that record (each triangle evaluates its own edges).
    const size_t NN = d->N;   // internal stride (incl. padding)
    std::vector<int> hr(3 * NN, -1);
    for (int p = 0; p < N; p++) {
        const int k = d->perm[p];
        const int4 v4 = d->hnb[p];
        const int v[3] = {v4.x, v4.y, v4.z};
        for (int i = 0; i < 3; i++) {
            int owner = 3 * k + i;
            if (v[i] >= 0) {
                const int n = d->perm[v[i] >> 2];
                if (n < k) owner = 3 * n + (v[i] & 3);
            }
            hr[(size_t)i * NN + p] = rank[owner];
        }
    }
    HIPCHK(d, hipStreamSynchronize(d->stream));
    drop_graph(d);
    if (d->rw) HIPCHK(d, hipFree(d->rw));
    if (d->rwd) HIPCHK(d, hipFree(d->rwd));
    d->rw = nullptr;
    d->rwd = nullptr;
    if (R == 0) return ANUGA_OK;
    HIPCHK(d, dalloc(&d->rw, 3 * NN));
    HIPCHK(d, dalloc(&d->rwd, (size_t)R));
    HIPCHK(d, hipMemcpy(d->rw, hr.data(), 3 * NN * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(d, hipMemcpy(d->rwd, rec.data(), R * sizeof(DeRiverwall), hipMemcpyHostToDevice));
    return ANUGA_OK;
}

static int ensure_iperm(anuga_domain *d)
{
    if (d->iperm_dev) return ANUGA_OK;
    const int N = d->ntri[0];
    std::vector<int> ip(N);
    for (int p = 0; p < N; p++) ip[d->perm[p]] = p;
    HIPCHK(d, dalloc(&d->iperm_dev, (size_t)N));
    HIPCHK(d, hipMemcpy(d->iperm_dev, ip.data(), N * sizeof(int), hipMemcpyHostToDevice));
    return ANUGA_OK;
}

static int ensure_out_f32(anuga_domain *d)
{
    if (d->out_f32) return ANUGA_OK;
    HIPCHK(d, dalloc(&d->out_f32, 3 * ((size_t)d->nodes + d->ntri[0])));
    return ANUGA_OK;
}

int anuga_set_output_nodes(anuga_domain *d, int64_t number_of_nodes,
                           const int64_t *number_of_triangles_per_node,
                           const int64_t *vertex_value_indices)
{
    if (!d) return ANUGA_E_ARG;
    int rc = single(d);
    if (rc) return rc;
    const int N = d->ntri[0];
    if (number_of_nodes < 1 || number_of_nodes >= (int64_t(1) << 30))
        return fail(d, ANUGA_E_ARG, "bad number_of_nodes");
    if (!number_of_triangles_per_node || !vertex_value_indices)
        return fail(d, ANUGA_E_ARG, "a node array is NULL");
    const int nodes = (int)number_of_nodes;
    std::vector<int> start(nodes + 1);
    int64_t total = 0;
    for (int j = 0; j < nodes; j++) {
        if (number_of_triangles_per_node[j] < 0)
            return fail(d, ANUGA_E_ARG, "negative number_of_triangles_per_node");
        start[j] = (int)total;
        total += number_of_triangles_per_node[j];
        if (total > 3 * (int64_t)N) break;
    }
    if (total != 3 * (int64_t)N)
        return fail(d, ANUGA_E_ARG, "number_of_triangles_per_node must sum to 3 N");
    start[nodes] = (int)total;
    std::vector<int> ip(N), tri(3 * (size_t)N);
    for (int p = 0; p < N; p++) ip[d->perm[p]] = p;
    for (int64_t i = 0; i < 3 * (int64_t)N; i++) {
        const int64_t v = vertex_value_indices[i];
        if (v < 0 || v >= 3 * (int64_t)N)
            return fail(d, ANUGA_E_ARG, "vertex_value_indices out of range");
        tri[i] = ip[v / 3];
    }
    HIPCHK(d, hipStreamSynchronize(d->stream));
    void *old[] = {d->node_start, d->node_tri, d->out_f32};
    for (void *q : old)
        if (q) HIPCHK(d, hipFree(q));
    d->node_start = d->node_tri = nullptr;
    d->out_f32 = nullptr;
    d->nodes = nodes;
    HIPCHK(d, dalloc(&d->node_start, (size_t)nodes + 1));
    HIPCHK(d, dalloc(&d->node_tri, 3 * (size_t)N));
    HIPCHK(d, hipMemcpy(d->node_start, start.data(), (nodes + 1) * sizeof(int),
                        hipMemcpyHostToDevice));
    HIPCHK(d, hipMemcpy(d->node_tri, tri.data(), 3 * (size_t)N * sizeof(int),
                        hipMemcpyHostToDevice));
    return ANUGA_OK;
}

int anuga_get_sww_frame(anuga_domain *d, double minimum_storable_height, float *node_values,
                        float *centroid_values)
{
    if (!d) return ANUGA_E_ARG;
    int rc = single(d);
    if (rc) return rc;
    if (node_values && !d->node_start)
        return fail(d, ANUGA_E_STATE, "no output nodes: call anuga_set_output_nodes first");
    const int N = d->ntri[0];
    if ((rc = ensure_iperm(d)) || (rc = ensure_out_f32(d))) return rc;
    const DeQ *S = d->substate ? d->B : d->A;
    float *on = d->out_f32, *oc = d->out_f32 + 3 * (size_t)d->nodes;
    if (node_values)
        hipLaunchKernelGGL(de_sww_nodes, dim3((d->nodes + DE_BLOCK - 1) / DE_BLOCK), dim3(DE_BLOCK),
                           0, d->stream, d->nodes, d->node_start, d->node_tri, S,
                           minimum_storable_height, on);
    if (centroid_values)
        hipLaunchKernelGGL(de_gather_f32, dim3((N + DE_BLOCK - 1) / DE_BLOCK), dim3(DE_BLOCK), 0,
                           d->stream, N, d->iperm_dev, (const double *)S, 4, 3,
                           make_int4(0, 1, 2, 0), oc);
    rc = launch_check(d);
    if (rc) return rc;
    if (node_values)
        HIPCHK(d, hipMemcpyAsync(node_values, on, 3 * (size_t)d->nodes * sizeof(float),
                                 hipMemcpyDeviceToHost, d->stream));
    if (centroid_values)
        HIPCHK(d, hipMemcpyAsync(centroid_values, oc, 3 * (size_t)N * sizeof(float),
                                 hipMemcpyDeviceToHost, d->stream));
    HIPCHK(d, hipStreamSynchronize(d->stream));
    return ANUGA_OK;
}

int anuga_get_static_output(anuga_domain *d, int quantity, float *node_values,
                            float *centroid_values)
{
    if (!d) return ANUGA_E_ARG;
    int rc = single(d);
    if (rc) return rc;
    if (quantity != ANUGA_ELEVATION && quantity != ANUGA_FRICTION)
        return fail(d, ANUGA_E_ARG, "static output is elevation or friction");
    if (node_values && !d->node_start)
        return fail(d, ANUGA_E_STATE, "no output nodes: call anuga_set_output_nodes first");
    const int N = d->ntri[0];
    if ((rc = ensure_iperm(d)) || (rc = ensure_out_f32(d))) return rc;
    const double *v = quantity == ANUGA_ELEVATION ? (const double *)d->A + 3 : d->eta;
    const int stride = quantity == ANUGA_ELEVATION ? 4 : 1;
    float *on = d->out_f32, *oc = d->out_f32 + 3 * (size_t)d->nodes;
    if (node_values)
        hipLaunchKernelGGL(de_node_values, dim3((d->nodes + DE_BLOCK - 1) / DE_BLOCK),
                           dim3(DE_BLOCK), 0, d->stream, d->nodes, d->node_start, d->node_tri, v,
                           stride, on);
    if (centroid_values)
        hipLaunchKernelGGL(de_gather_f32, dim3((N + DE_BLOCK - 1) / DE_BLOCK), dim3(DE_BLOCK), 0,
                           d->stream, N, d->iperm_dev, v, stride, 1, make_int4(0, 0, 0, 0), oc);
    rc = launch_check(d);
    if (rc) return rc;
    if (node_values)
        HIPCHK(d, hipMemcpyAsync(node_values, on, (size_t)d->nodes * sizeof(float),
                                 hipMemcpyDeviceToHost, d->stream));
    if (centroid_values)
        HIPCHK(d, hipMemcpyAsync(centroid_values, oc, (size_t)N * sizeof(float),
                                 hipMemcpyDeviceToHost, d->stream));
    HIPCHK(d, hipStreamSynchronize(d->stream));
    return ANUGA_OK;
}

int anuga_set_gauges(anuga_domain *d, int64_t ngauges, const int64_t *triangles,
                     int64_t capacity)
{
    if (!d) return ANUGA_E_ARG;
    int rc = single(d);
    if (rc) return rc;
    const int N = d->ntri[0];
    if (ngauges < 0 || ngauges > N) return fail(d, ANUGA_E_ARG, "bad gauge count");
    if (ngauges > 0 && (!triangles || capacity < 1 || capacity > (int64_t(1) << 24) ||
                        capacity * ngauges > (int64_t(1) << 31)))
        return fail(d, ANUGA_E_ARG, "gauge triangles NULL or bad capacity");
    std::vector<int> ip(N), gt(ngauges);
    for (int p = 0; p < N; p++) ip[d->perm[p]] = p;
    for (int64_t g = 0; g < ngauges; g++) {
        if (triangles[g] < 0 || triangles[g] >= N)
            return fail(d, ANUGA_E_ARG, "gauge triangle out of range");
        gt[g] = ip[triangles[g]];
    }
    HIPCHK(d, hipStreamSynchronize(d->stream));
    drop_graph(d);
    void *old[] = {d->gtri, d->gtimes, d->gvals};
    for (void *q : old)
        if (q) HIPCHK(d, hipFree(q));
    d->gtri = nullptr;
    d->gtimes = d->gvals = nullptr;
    d->G = d->gcap = 0;
    d->grows_host = 0;
    if (ngauges > 0) {
        HIPCHK(d, dalloc(&d->gtri, (size_t)ngauges));
        HIPCHK(d, dalloc(&d->gtimes, (size_t)capacity));
        HIPCHK(d, dalloc(&d->gvals, (size_t)capacity * ngauges * 4));
        HIPCHK(d, hipMemcpy(d->gtri, gt.data(), ngauges * sizeof(int), hipMemcpyHostToDevice));
        d->G = (int)ngauges;
        d->gcap = (int)capacity;
    }
    hipLaunchKernelGGL(de_gauge_control, dim3(1), dim3(64), 0, d->stream, d->ctl, 0, d->gcap);
    return launch_check(d);
}

int anuga_sample_gauges(anuga_domain *d, double relative_time)
{
    if (!d) return ANUGA_E_ARG;
    int rc = single(d);
    if (rc) return rc;
    if (!d->G) return fail(d, ANUGA_E_STATE, "no gauges set");
    if (d->grows_host >= d->gcap)
        return fail(d, ANUGA_E_STATE, "gauge rows full: download them first");
    hipLaunchKernelGGL(de_sample_gauges<true>, dim3(1), dim3(256), 0, d->stream, d->ctl, d->G,
                       d->gtri, d->substate ? d->B : d->A, relative_time, d->gtimes, d->gvals);
    d->grows_host++;
    return launch_check(d);
}

int anuga_get_gauge_samples(anuga_domain *d, int64_t max_rows, int64_t *nrows, double *times,
                            double *values)
{
    if (!d || !nrows) return ANUGA_E_ARG;
    int rc = single(d);
    if (rc) return rc;
    *nrows = 0;
    if (!d->G) return fail(d, ANUGA_E_STATE, "no gauges set");
    rc = sync_scalars(d);
    if (rc) return rc;
    const int rows = d->hctl[0].grows;
    if (rows > max_rows) return fail(d, ANUGA_E_ARG, "max_rows is below the rows recorded");
    if (rows > 0 && (!times || !values)) return fail(d, ANUGA_E_ARG, "times or values is NULL");
    if (rows > 0) {
        HIPCHK(d, hipMemcpy(times, d->gtimes, rows * sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(d, hipMemcpy(values, d->gvals, (size_t)rows * d->G * 4 * sizeof(double),
                            hipMemcpyDeviceToHost));
    }
    hipLaunchKernelGGL(de_gauge_control, dim3(1), dim3(64), 0, d->stream, d->ctl, 1, 0);
    rc = launch_check(d);
    if (rc) return rc;
    HIPCHK(d, hipStreamSynchronize(d->stream));
    d->grows_host = 0;
    *nrows = rows;
    return ANUGA_OK;
}

int anuga_distribute_to_vertices_and_edges(anuga_domain *d, double *mass_error)
{
    if (!d) return ANUGA_E_ARG;
    int rc = single(d);
    if (rc) return rc;
    rc = distribute(d);
    if (rc) return rc;
    if (mass_error) *mass_error = d->hctl->mass;
    return ANUGA_OK;
}

static int lts_fluxes(anuga_domain *d, double *flux_timestep);
static int op_kend(const anuga_domain *d);

int anuga_compute_fluxes(anuga_domain *d, double *flux_timestep)
{
    if (!d) return ANUGA_E_ARG;
    int rc = single(d);
    if (rc) return rc;
    if (d->substate != 0) return fail(d, ANUGA_E_STATE, "compute_fluxes called inside an rk2 step");
    rc = ready_to_step(d);
    if (rc) return rc;
    ensure_live(d);
    if (d->lts_M > 1) {
        launch_extrapolate(d, d->A);
        rc = lts_fluxes(d, flux_timestep);
        if (d->timing) timer_flush(d);
        return rc;
    }
    if (fused(d)) {
        launch_xflux(d, 0, d->A);
    } else {
        launch_extrapolate(d, d->A);
        launch_flux(d, 0, d->A);
    }
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
    int rc = single(d);
    if (rc) return rc;
    if (d->substate != 0) return fail(d, ANUGA_E_STATE, "update called twice in a step");
    ensure_live(d);
    const bool rk = d->P.timestep_fluxcalls > 1;
    hipLaunchKernelGGL(de_set_dt, dim3(1), dim3(64), 0, d->stream, d->ctl, dt);
    launch_update(d, d->A, rk ? d->B : d->A);
    if (!rk) hipLaunchKernelGGL(de_integrate, dim3(1), dim3(64), 0, d->stream, d->ctl, 1);
    rc = launch_check(d);
    if (rc) return rc;
    if (rk) d->substate = 1;
    return num_negative ? read_negative(d, num_negative) : ANUGA_OK;
}

// the substeps after the first of an rk2 / rk3 step (same dt)
static int later_substeps(anuga_domain *d, double dt, int64_t *num_negative, int method)
{
    int rc = single(d);
    if (rc) return rc;
    if (d->P.timestep_fluxcalls != method)
        return fail(d, ANUGA_E_STATE, "substeps of another timestepping method");
    if (d->substate != 1) return fail(d, ANUGA_E_STATE, "second substep without a first");
    hipLaunchKernelGGL(de_set_dt, dim3(1), dim3(64), 0, d->stream, d->ctl, dt);
    enqueue_later_substeps(d);
    hipLaunchKernelGGL(de_integrate, dim3(1), dim3(64), 0, d->stream, d->ctl, method);
    rc = launch_check(d);
    if (rc) return rc;
    d->substate = 0;
    return num_negative ? read_negative(d, num_negative) : ANUGA_OK;
}

int anuga_rk2_second_substep(anuga_domain *d, double dt, int64_t *num_negative)
{
    if (!d) return ANUGA_E_ARG;
    return later_substeps(d, dt, num_negative, 2);
}

int anuga_rk3_later_substeps(anuga_domain *d, double dt, int64_t *num_negative)
{
    if (!d) return ANUGA_E_ARG;
    return later_substeps(d, dt, num_negative, 3);
}

// ---------------------------------------------------------------------------
// local flux updating (sw_domain_orig.c:509-669 and the skip at :829)
// ---------------------------------------------------------------------------
// The flux call of a domain with local flux updating (euler: substep 0):
// the edges due this call, the stored terms of the others; the CFL timestep
// is the reference's static one, restarted from 1e100 only when the step
// count allows the timestep to grow (:813-817).
static int lts_fluxes(anuga_domain *d, double *flux_timestep)
{
    if (d->lts_call == INT32_MAX)
        return fail(d, ANUGA_E_STATE, "local flux updating: too many flux calls");
    d->lts_call++;
    DeBoundary bc{d->bkind, d->bval, d->M, d->bsstride};
    hipLaunchKernelGGL(de_flux_lts, dim3(d->nblocks), dim3(DE_BLOCK), 0, d->stream, dev_mesh(d),
                       dev_params(d->P), d->A, (const double *)d->E, bc, d->flags, d->ex[0],
                       d->ex[1], d->ex[2], d->lts, d->lts_call, d->ctl, d->part,
                       d->part + d->nblocks);
    launch_reduce(d, 0);
    int rc = launch_check(d);
    if (rc) return rc;
    rc = sync_scalars(d);
    if (rc) return rc;
    const double fresh = d->hctl->flux_dt;
    d->lts_dt = d->lts_allow == 1 ? fresh : fmin(d->lts_dt, fresh);
    if (flux_timestep) *flux_timestep = d->lts_dt;
    return ANUGA_OK;
}

int anuga_set_max_flux_update_frequency(anuga_domain *d, int64_t max_flux_update_frequency)
{
    if (!d) return ANUGA_E_ARG;
    int rc = single(d);
    if (rc) return rc;
    const int64_t M = max_flux_update_frequency;
    if (M < 1 || M > (int64_t(1) << 30) || (M & (M - 1)) != 0)
        return fail(d, ANUGA_E_ARG, "max_flux_update_frequency must be a power of 2 (at most 2^30)");
    if (M == 1) {
        d->lts_M = 1;
        return ANUGA_OK;
    }
    if (d->P.timestep_fluxcalls != 1)
        return fail(d, ANUGA_E_STATE,
                    "Local extrapolation and flux updating only supported with euler timestepping");
    if (d->mixed || d->rw || d->P.low_froude || d->P.use_sloped_mannings)
        return fail(d, ANUGA_E_STATE, "local flux updating runs the FP64 kernels without "
                                      "riverwalls, low_froude or sloped friction");
    const size_t n3 = 3 * (size_t)d->N;
    HIPCHK(d, hipStreamSynchronize(d->stream));
    if (!d->lts.e0) {
        HIPCHK(d, dalloc(&d->lts.e0, n3));
        HIPCHK(d, dalloc(&d->lts.e1, n3));
        HIPCHK(d, dalloc(&d->lts.e2, n3));
        HIPCHK(d, dalloc(&d->lts.pg, n3));
        HIPCHK(d, dalloc(&d->lts.et, n3));
        HIPCHK(d, dalloc(&d->lts.fuf, n3));
        HIPCHK(d, dalloc(&d->lts.acf, n3));
        HIPCHK(d, dalloc(&d->lts.tmin, (size_t)d->N));
        HIPCHK(d, dalloc(&d->lts.unf, n3));
    }
    hipLaunchKernelGGL(de_lts_init, dim3((unsigned)((n3 + DE_BLOCK - 1) / DE_BLOCK)),
                       dim3(DE_BLOCK), 0, d->stream, (int)n3, d->lts);
    d->lts_M = (int)M;
    d->lts_call = 0;
    d->lts_cyc = -1;
    d->lts_dt = 0.0;
    d->lts_allow = 1;
    return launch_check(d);
}

int anuga_compute_flux_update_frequency(anuga_domain *d, double timestep)
{
    if (!d) return ANUGA_E_ARG;
    int rc = single(d);
    if (rc) return rc;
    if (d->lts_M == 1) return ANUGA_OK;
    ensure_live(d);
    d->lts_cyc++;
    if (d->lts_cyc == d->lts_M) d->lts_cyc = 0;
    const int kend = op_kend(d);
    const dim3 grid((kend + DE_BLOCK - 1) / DE_BLOCK);
    hipLaunchKernelGGL(de_lts_part1, grid, dim3(DE_BLOCK), 0, d->stream, kend, d->N, d->nb,
                       d->lts, d->lts_cyc, d->lts_M, timestep);
    hipLaunchKernelGGL(de_lts_part2, grid, dim3(DE_BLOCK), 0, d->stream, kend, d->N, d->nb,
                       d->lts, d->lts_cyc + 1);
    d->lts_allow = d->lts_cyc + 1 == d->lts_M ? 1 : 0;
    return launch_check(d);
}

// ---------------------------------------------------------------------------
// per-operation calls in the reference's order
// ---------------------------------------------------------------------------
static int op_ready(anuga_domain *d)
{
    int rc = single(d);
    if (rc) return rc;
    if (d->substate != 0)
        return fail(d, ANUGA_E_STATE, "per-operation call inside a fused rk step");
    rc = ready_to_step(d);
    if (rc) return rc;
    ensure_live(d);
    return ANUGA_OK;
}

static int op_kend(const anuga_domain *d)
{
    return d->kbeg[0] + d->ntri[0];
}

static dim3 op_grid(const anuga_domain *d)
{
    return dim3((op_kend(d) + DE_BLOCK - 1) / DE_BLOCK);
}

int anuga_backup_conserved_quantities(anuga_domain *d)
{
    if (!d) return ANUGA_E_ARG;
    int rc = op_ready(d);
    if (rc) return rc;
    hipLaunchKernelGGL(de_backup, op_grid(d), dim3(DE_BLOCK), 0, d->stream, op_kend(d), d->A,
                       d->B);
    rc = launch_check(d);
    if (rc) return rc;
    d->backup_valid = true;
    return ANUGA_OK;
}

int anuga_saxpy_conserved_quantities(anuga_domain *d, double a, double b)
{
    if (!d) return ANUGA_E_ARG;
    int rc = op_ready(d);
    if (rc) return rc;
    if (!d->backup_valid)
        return fail(d, ANUGA_E_STATE, "saxpy_conserved_quantities without a backup");
    hipLaunchKernelGGL(de_saxpy, op_grid(d), dim3(DE_BLOCK), 0, d->stream, op_kend(d), a, b,
                       d->A, d->B);
    return launch_check(d);
}

int anuga_divide_conserved_quantities(anuga_domain *d, double divisor)
{
    if (!d) return ANUGA_E_ARG;
    int rc = op_ready(d);
    if (rc) return rc;
    hipLaunchKernelGGL(de_divide, op_grid(d), dim3(DE_BLOCK), 0, d->stream, op_kend(d), divisor,
                       d->A);
    return launch_check(d);
}

int anuga_update_boundary(anuga_domain *d)
{
    if (!d) return ANUGA_E_ARG;
    return op_ready(d);
}

int anuga_compute_fluxes_ext_central(anuga_domain *d, double timestep, double *flux_timestep)
{
    if (!d || !flux_timestep) return ANUGA_E_ARG;
    int rc = op_ready(d);
    if (rc) return rc;
    if (d->e_stale)
        return fail(d, ANUGA_E_STATE,
                    "compute_fluxes needs the edge values of distribute_to_vertices_and_edges");
    if (d->lts_M > 1) {
        rc = lts_fluxes(d, flux_timestep);
        if (rc) return rc;
        if (d->timing) timer_flush(d);
        d->op_flux_calls++;
        return ANUGA_OK;
    }
    const int sub = (int)(d->op_flux_calls % d->P.timestep_fluxcalls);
    launch_flux(d, 0, d->A);
    launch_reduce(d, sub == 0 ? 0 : (sub == 1 ? 1 : 4));
    rc = launch_check(d);
    if (rc) return rc;
    rc = sync_scalars(d);
    if (rc) return rc;
    if (d->timing) timer_flush(d);
    d->op_flux_calls++;
    *flux_timestep = sub == 0 ? d->hctl->flux_dt : timestep;
    return ANUGA_OK;
}

int anuga_compute_forcing_terms(anuga_domain *d)
{
    if (!d) return ANUGA_E_ARG;
    int rc = op_ready(d);
    if (rc) return rc;
    const size_t N = (size_t)d->N;
    for (int i = 0; i < 2; i++)
        if (!d->si[i]) {
            HIPCHK(d, hipStreamSynchronize(d->stream));
            HIPCHK(d, dalloc(&d->si[i], N));
            HIPCHK(d, hipMemsetAsync(d->si[i], 0, N * sizeof(double), d->stream));
        }
    const int kend = op_kend(d);
    if (d->P.use_sloped_mannings && !d->mixed) {
        if (d->e_stale)
            return fail(d, ANUGA_E_STATE, "sloped friction needs the edge values of "
                                          "distribute_to_vertices_and_edges");
        hipLaunchKernelGGL(de_friction<true>, op_grid(d), dim3(DE_BLOCK), 0, d->stream, kend,
                           dev_params(d->P), d->A, d->eta, d->si[0], d->si[1],
                           (const double *)d->E, d->N, d->vx);
    } else {
        hipLaunchKernelGGL(de_friction<false>, op_grid(d), dim3(DE_BLOCK), 0, d->stream, kend,
                           dev_params(d->P), d->A, d->eta, d->si[0], d->si[1], nullptr, 0,
                           nullptr);
    }
    return launch_check(d);
}

int anuga_update_centroids(anuga_domain *d, double dt)
{
    if (!d) return ANUGA_E_ARG;
    int rc = op_ready(d);
    if (rc) return rc;
    const size_t N = (size_t)d->N;
    for (int i = 0; i < 2; i++)
        if (!d->si[i]) {   // no forcing terms yet: zero semi-implicit updates
            HIPCHK(d, hipStreamSynchronize(d->stream));
            HIPCHK(d, dalloc(&d->si[i], N));
            HIPCHK(d, hipMemsetAsync(d->si[i], 0, N * sizeof(double), d->stream));
        }
    hipLaunchKernelGGL(de_set_dt, dim3(1), dim3(64), 0, d->stream, d->ctl, dt);
    hipLaunchKernelGGL(de_update_centroids, op_grid(d), dim3(DE_BLOCK), 0, d->stream, op_kend(d),
                       d->ctl, d->A, d->ex[0], d->ex[1], d->ex[2], d->si[0], d->si[1]);
    d->e_stale = true;
    rc = launch_check(d);
    return rc ? rc : sync_scalars(d);
}

int anuga_fix_negative_cells(anuga_domain *d, int64_t *num_negative)
{
    if (!d) return ANUGA_E_ARG;
    int rc = op_ready(d);
    if (rc) return rc;
    hipLaunchKernelGGL(de_fix_negative, op_grid(d), dim3(DE_BLOCK), 0, d->stream, op_kend(d),
                       d->A, d->dneg);
    rc = launch_check(d);
    if (rc) return rc;
    int64_t n = 0;
    rc = read_negative(d, &n);
    if (rc) return rc;
    if (num_negative) *num_negative = n;
    return ANUGA_OK;
}

int anuga_boundary_flux_integral_operator(anuga_domain *d, double timestep)
{
    if (!d) return ANUGA_E_ARG;
    int rc = op_ready(d);
    if (rc) return rc;
    hipLaunchKernelGGL(de_set_dt, dim3(1), dim3(64), 0, d->stream, d->ctl, timestep);
    hipLaunchKernelGGL(de_integrate, dim3(1), dim3(64), 0, d->stream, d->ctl,
                       (int)d->P.timestep_fluxcalls);
    return launch_check(d);
}

int anuga_evolve_steps(anuga_domain *d, int64_t max_steps, const anuga_time_control *tc,
                       anuga_step_state *st, double *dt_history)
{
    if (!d || !tc || !st || max_steps < 0) return ANUGA_E_ARG;
    int rc = single(d);
    return rc ? rc : evolve_steps(d, max_steps, tc, st, dt_history, false);
}

int anuga_get_boundary_flux_sums(anuga_domain *d, double *sums)
{
    if (!d || !sums) return ANUGA_E_ARG;
    int rc = single(d);
    if (rc) return rc;
    rc = sync_scalars(d);
    if (rc) return rc;
    sums[0] = d->hctl->bflux0;
    sums[1] = d->P.timestep_fluxcalls >= 2 ? d->hctl->bflux1 : 0.0;
    sums[2] = d->P.timestep_fluxcalls == 3 ? d->hctl->bflux2 : 0.0;
    return ANUGA_OK;
}

int anuga_get_boundary_flux_integral(anuga_domain *d, double *integral)
{
    if (!d || !integral) return ANUGA_E_ARG;
    int rc = single(d);
    if (rc) return rc;
    rc = sync_scalars(d);
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
    int rc = single(d);
    if (rc) return rc;
    rc = volumes(d);
    if (rc) return rc;
    *volume = d->hctl->volume;
    return ANUGA_OK;
}

// the current device's scratch address (looked up once per device)
static int scratch_address(double **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(nullptr, ANUGA_E_HIP, hipGetErrorString(e));
    if (dev < 0 || dev >= 256) return fail(nullptr, ANUGA_E_HIP, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(nullptr, ANUGA_E_HIP, std::string("scratch address: ") +
                                                  hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return ANUGA_OK;
}

int anuga_water_volume_noalloc(int64_t N, const double *stage, const double *elevation,
                               const double *area, double *volume, void *stream)
{
    if (N <= 0 || !stage || !elevation || !area || !volume)
        return fail(nullptr, ANUGA_E_ARG, "water_volume_noalloc: bad arguments");
    if (N > kNoAllocMaxN)
        return fail(nullptr, ANUGA_E_ARG, "water_volume_noalloc: at most 2^27 triangles");
    double *sc = nullptr;
    int rc = scratch_address(&sc);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    DeControl *ctl = reinterpret_cast<DeControl *>(sc);
    double *chunks = sc + kNoAllocCtl, *part = chunks + kNoAllocChunks;
    const int n = (int)N, nb = (int)((N + DE_BLOCK - 1) / DE_BLOCK);
    // the reduction of anuga_get_water_volume (de_volume, MODE 3), over the
    // caller's arrays
    hipLaunchKernelGGL(de_control_init, dim3(1), dim3(64), 0, st, ctl, n, nb);
    hipLaunchKernelGGL(de_volume_arrays, dim3(nb), dim3(DE_BLOCK), 0, st, n, stage, elevation,
                       area, part);
    hipLaunchKernelGGL(de_reduce_chunks<3>, dim3(DE_RCHUNKS, 1), dim3(256), 0, st, part, part,
                       part, ctl, chunks);
    hipLaunchKernelGGL(de_reduce_final<3>, dim3(1), dim3(64), 0, st, chunks, ctl);
    hipLaunchKernelGGL(de_copy_volume, dim3(1), dim3(64), 0, st, ctl, volume);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(nullptr, ANUGA_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
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

int anuga_build_config(int64_t *block, int64_t *mixed_waves)
{
    if (!block || !mixed_waves) return ANUGA_E_ARG;
    *block = DE_BLOCK;
    *mixed_waves = DE_MIXED_WAVES;
    return ANUGA_OK;
}

int anuga_tile_order(int64_t N, const double *centroids, const int64_t *neighbours,
                     int64_t *order)
{
    if (N < 0 || N >= (int64_t(1) << 29) || (N > 0 && (!centroids || !order))) return ANUGA_E_ARG;
    if (neighbours)
        for (int64_t i = 0; i < 3 * N; i++)
            if (neighbours[i] >= N) return ANUGA_E_ARG;
    const std::vector<int> o = tile_order((int)N, centroids, neighbours);
    for (int64_t k = 0; k < N; k++) order[k] = o[k];
    return ANUGA_OK;
}

int anuga_tile_edges(int64_t N, const int32_t *nb, int64_t *estart, int64_t *esplit,
                     int64_t edge_cap, uint32_t *edges, int64_t *nedges)
{
    if (N < 0 || N >= (int64_t(1) << 29) || (N > 0 && !nb) || !estart || !esplit || !nedges)
        return ANUGA_E_ARG;
    std::vector<int4> hn(N);
    for (int64_t k = 0; k < N; k++) hn[k] = make_int4(nb[3 * k], nb[3 * k + 1], nb[3 * k + 2], 0);
    std::vector<int> st, h, est, esp;
    std::vector<unsigned> sl, e;
    if (tile_halos((int)N, hn.data(), st, h, sl) < 0) return ANUGA_E_ARG;
    tile_edges((int)N, hn.data(), sl, est, esp, e);
    *nedges = (int64_t)e.size();
    if ((int64_t)e.size() > edge_cap || (!edges && !e.empty())) return ANUGA_E_ARG;
    for (size_t b = 0; b < est.size(); b++) estart[b] = est[b];
    for (size_t b = 0; b < esp.size(); b++) esplit[b] = esp[b];
    for (size_t q = 0; q < e.size(); q++) edges[q] = e[q];
    return ANUGA_OK;
}

int anuga_tile_halos(int64_t N, const int32_t *nb, int64_t *start, int64_t halo_cap,
                     int32_t *halo, uint32_t *slot, int64_t *nhalo, int64_t *hcap)
{
    if (N < 0 || N >= (int64_t(1) << 29) || (N > 0 && (!nb || !slot)) || !start || !nhalo || !hcap)
        return ANUGA_E_ARG;
    std::vector<int4> hn(N);
    for (int64_t k = 0; k < N; k++) hn[k] = make_int4(nb[3 * k], nb[3 * k + 1], nb[3 * k + 2], 0);
    std::vector<int> st, h;
    std::vector<unsigned> sl;
    const int c = tile_halos((int)N, hn.data(), st, h, sl);
    if (c < 0) return ANUGA_E_ARG;
    *nhalo = (int64_t)h.size();
    *hcap = c;
    if ((int64_t)h.size() > halo_cap || (!halo && !h.empty())) return ANUGA_E_ARG;
    for (size_t b = 0; b < st.size(); b++) start[b] = st[b];
    for (size_t q = 0; q < h.size(); q++) halo[q] = h[q];
    for (int64_t k = 0; k < N; k++) slot[k] = sl[k];
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

// ---------------------------------------------------------------------------
// ensembles
// ---------------------------------------------------------------------------
int anuga_ensemble_create(int64_t nscenarios, const int64_t *number_of_elements,
                          const int64_t *boundary_lengths, const anuga_params *params,
                          anuga_domain **out)
{
    if (!out) return fail(nullptr, ANUGA_E_ARG, "out is NULL");
    *out = nullptr;
    if (!number_of_elements || !boundary_lengths)
        return fail(nullptr, ANUGA_E_ARG, "size arrays are NULL");
    return domain_alloc(params, nscenarios, number_of_elements, boundary_lengths, out);
}

int anuga_ensemble_set_mesh(anuga_domain *d, int64_t scenario, const anuga_mesh *mesh)
{
    if (!d) return ANUGA_E_ARG;
    if (scenario < 0 || scenario >= d->S) return fail(d, ANUGA_E_ARG, "scenario index out of range");
    if (d->mesh_set[scenario]) return fail(d, ANUGA_E_STATE, "the scenario's mesh is already set");
    int rc = validate_mesh(d, mesh, d->ntri[scenario], d->nbnd[scenario]);
    return rc ? rc : upload_mesh(d, (int)scenario, mesh);
}

int anuga_ensemble_set_centroid_values(anuga_domain *d, int64_t scenario, int quantity,
                                       const double *values)
{
    if (!d) return ANUGA_E_ARG;
    int rc = scenario_ok(d, scenario);
    return rc ? rc : set_centroids(d, (int)scenario, quantity, values);
}

int anuga_ensemble_get_centroid_values(anuga_domain *d, int64_t scenario, int quantity,
                                       double *values)
{
    if (!d) return ANUGA_E_ARG;
    int rc = scenario_ok(d, scenario);
    return rc ? rc : get_centroids(d, (int)scenario, quantity, values);
}

int anuga_ensemble_set_boundary(anuga_domain *d, int64_t scenario, const int32_t *kind,
                                const double *dirichlet_values)
{
    if (!d) return ANUGA_E_ARG;
    int rc = scenario_ok(d, scenario);
    return rc ? rc : set_boundary(d, (int)scenario, kind, dirichlet_values);
}

int anuga_ensemble_evolve_steps(anuga_domain *d, int64_t max_steps,
                                const anuga_time_control *tc, anuga_step_state *states)
{
    if (!d || !tc || !states || max_steps < 0) return ANUGA_E_ARG;
    return evolve_steps(d, max_steps, tc, states, nullptr, true);
}

int anuga_ensemble_distribute_to_vertices_and_edges(anuga_domain *d, double *mass_errors)
{
    if (!d) return ANUGA_E_ARG;
    int rc = distribute(d);
    if (rc) return rc;
    if (mass_errors)
        for (int s = 0; s < d->S; s++) mass_errors[s] = d->hctl[s].mass;
    return ANUGA_OK;
}

int anuga_ensemble_get_water_volumes(anuga_domain *d, double *volumes_out)
{
    if (!d || !volumes_out) return ANUGA_E_ARG;
    int rc = ready_to_step(d);
    if (rc) return rc;
    rc = volumes(d);
    if (rc) return rc;
    for (int s = 0; s < d->S; s++) volumes_out[s] = d->hctl[s].volume;
    return ANUGA_OK;
}

int anuga_ensemble_get_boundary_flux_integrals(anuga_domain *d, double *integrals)
{
    if (!d || !integrals) return ANUGA_E_ARG;
    int rc = sync_scalars(d);
    if (rc) return rc;
    for (int s = 0; s < d->S; s++) integrals[s] = d->hctl[s].integral;
    return ANUGA_OK;
}

int anuga_ensemble_size(const anuga_domain *d, int64_t *nscenarios)
{
    if (!d || !nscenarios) return ANUGA_E_ARG;
    *nscenarios = d->S;
    return ANUGA_OK;
}

// ---------------------------------------------------------------------------
// gradients
// ---------------------------------------------------------------------------
int anuga_adjoint_record(anuga_domain *d, int64_t nsteps, double dt, int64_t checkpoint_every,
                         int64_t ngauges, const int64_t *gauge_triangles, int64_t obs_every,
                         double *gauge_stage)
{
    if (!d) return ANUGA_E_ARG;
    int rc = single(d);
    if (rc) return rc;
    if (d->mixed) return fail(d, ANUGA_E_ARG, "gradients need the FP64 path");
    if (d->rw) return fail(d, ANUGA_E_ARG, "gradients of runs with riverwalls are not supported");
    if (d->P.low_froude || d->P.use_sloped_mannings)
        return fail(d, ANUGA_E_ARG, "gradients need low_froude 0 and flat friction");
    for (signed char kb : d->hbkind)
        if (kb > ANUGA_BC_TRANSMISSIVE_CENTROID)
            return fail(d, ANUGA_E_ARG, "gradients support reflective, transmissive and "
                                        "Dirichlet boundaries only");
    if (d->bsstride) return fail(d, ANUGA_E_ARG, "gradients need constant boundary values");
    if (d->P.timestep_fluxcalls != 2) return fail(d, ANUGA_E_ARG, "gradients need rk2 (DE1)");
    if (nsteps <= 0 || checkpoint_every <= 0 || obs_every <= 0 || !(dt > 0.0) || ngauges < 0 ||
        (ngauges > 0 && (!gauge_triangles || !gauge_stage)))
        return fail(d, ANUGA_E_ARG, "bad run description");
    if (d->substate != 0) return fail(d, ANUGA_E_STATE, "inside an rk2 step");
    rc = ready_to_step(d);
    if (rc) return rc;
    const int N = d->N, G = (int)ngauges;
    std::vector<int> iperm(N), g(std::max(G, 1));
    for (int p = 0; p < N; p++) iperm[d->perm[p]] = p;
    std::vector<char> seen(N, 0);
    for (int i = 0; i < G; i++) {
        const int64_t t = gauge_triangles[i];
        if (t < 0 || t >= N) return fail(d, ANUGA_E_ARG, "gauge triangle out of range");
        if (seen[t]) return fail(d, ANUGA_E_ARG, "gauge triangles must be distinct");
        seen[t] = 1;
        g[i] = iperm[t];
    }
    delete d->adj;
    d->adj = nullptr;
    AdjRun *a = new AdjRun();
    a->nsteps = nsteps;
    a->every = std::min(checkpoint_every, nsteps);
    a->obs_every = obs_every;
    a->nckpt = (nsteps + a->every - 1) / a->every;
    a->nobs = nsteps / obs_every;
    a->G = G;
    a->dt = dt;
    const size_t n3 = 3 * (size_t)N;
#define ACHK(call)                                                                  \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                     \
            delete a;                                                               \
            return fail(d, ANUGA_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
        }                                                                           \
    } while (0)
    ACHK(dalloc(&a->ckpt, a->nckpt * (size_t)N));
    ACHK(dalloc(&a->seg, a->every * (size_t)N));
    ACHK(dalloc(&a->S1, N));
    ACHK(dalloc(&a->E0, 4 * n3));
    ACHK(dalloc(&a->E1, 4 * n3));
    ACHK(dalloc(&a->ex0, n3));
    ACHK(dalloc(&a->ex1, n3));
    ACHK(dalloc(&a->f0, N));
    ACHK(dalloc(&a->f1, N));
    ACHK(dalloc(&a->lam, n3));
    ACHK(dalloc(&a->lamn, n3));
    ACHK(dalloc(&a->lam1, n3));
    ACHK(dalloc(&a->lamE, 4 * n3));
    ACHK(dalloc(&a->lamX, 16 * (size_t)N));
    ACHK(dalloc(&a->lam_ex, n3));
    ACHK(dalloc(&a->zbar, N));
    ACHK(dalloc(&a->nbar, N));
    ACHK(dalloc(&a->rec, std::max<int64_t>(a->nobs * G, 1)));
    ACHK(dalloc(&a->gauges, std::max(G, 1)));
    ACHK(dalloc(&a->neg, 1));
    ACHK(hipMemcpy(a->gauges, g.data(), std::max(G, 1) * sizeof(int), hipMemcpyHostToDevice));
#undef ACHK
    d->adj = a;
    ensure_live(d);
    hipLaunchKernelGGL(de_set_dt, dim3(1), dim3(64), 0, d->stream, d->ctl, dt);
    HIPCHK(d, hipMemcpyAsync(a->ckpt, d->A, N * sizeof(DeQ), hipMemcpyDeviceToDevice, d->stream));
    for (int64_t s = 1; s <= nsteps; s++) {
        step_fixed(d, d->A);
        if (s % a->every == 0 && s < nsteps)
            HIPCHK(d, hipMemcpyAsync(a->ckpt + (s / a->every) * (size_t)N, d->A, N * sizeof(DeQ),
                                     hipMemcpyDeviceToDevice, d->stream));
        if (G > 0 && s % obs_every == 0)
            hipLaunchKernelGGL(adj_k_gauges, dim3((G + 63) / 64), dim3(64), 0, d->stream, G,
                               a->gauges, d->A, a->rec + (s / obs_every - 1) * (size_t)G);
    }
    rc = launch_check(d);
    if (rc) return rc;
    rc = sync_scalars(d);
    if (rc) return rc;
    if (G > 0)
        HIPCHK(d, hipMemcpy(gauge_stage, a->rec, a->nobs * G * sizeof(double),
                            hipMemcpyDeviceToHost));
    return ANUGA_OK;
}

int anuga_adjoint_gradient(anuga_domain *d, const double *dloss_dstage, double *grad_friction,
                           double *grad_elevation, double *grad_state0)
{
    if (!d) return ANUGA_E_ARG;
    AdjRun *a = d->adj;
    if (!a) return fail(d, ANUGA_E_STATE, "no recorded run: call anuga_adjoint_record first");
    if (a->G > 0 && !dloss_dstage) return fail(d, ANUGA_E_ARG, "dloss_dstage is NULL");
    const int N = d->N;
    const size_t n3 = 3 * (size_t)N;
    if (a->G > 0)
        HIPCHK(d, hipMemcpy(a->rec, dloss_dstage, a->nobs * a->G * sizeof(double),
                            hipMemcpyHostToDevice));
    HIPCHK(d, hipMemsetAsync(a->lam, 0, n3 * sizeof(double), d->stream));
    HIPCHK(d, hipMemsetAsync(a->zbar, 0, N * sizeof(double), d->stream));
    HIPCHK(d, hipMemsetAsync(a->nbar, 0, N * sizeof(double), d->stream));
    ensure_live(d);
    hipLaunchKernelGGL(de_set_dt, dim3(1), dim3(64), 0, d->stream, d->ctl, a->dt);
    for (int64_t j = a->nckpt - 1; j >= 0; j--) {
        const int64_t s0 = j * a->every, s1 = std::min(s0 + a->every, a->nsteps);
        HIPCHK(d, hipMemcpyAsync(a->seg, a->ckpt + j * (size_t)N, N * sizeof(DeQ),
                                 hipMemcpyDeviceToDevice, d->stream));
        for (int64_t t = 1; t < s1 - s0; t++) {
            DeQ *x = a->seg + t * (size_t)N;
            HIPCHK(d, hipMemcpyAsync(x, x - N, N * sizeof(DeQ), hipMemcpyDeviceToDevice, d->stream));
            step_fixed(d, x);
        }
        for (int64_t s = s1; s > s0; s--) {
            if (a->G > 0 && s % a->obs_every == 0)
                hipLaunchKernelGGL(adj_k_gauge_seed, dim3((a->G + 63) / 64), dim3(64), 0, d->stream,
                                   a->G, a->gauges, a->rec + (s / a->obs_every - 1) * (size_t)a->G,
                                   a->lam);
            int rc = adj_step(d, a, a->seg + (s - 1 - s0) * (size_t)N);
            if (rc) return rc;
        }
    }
    HIPCHK(d, hipStreamSynchronize(d->stream));
    std::vector<double> h(n3);
    auto unpermute = [&](const double *src, double *dst) {
        for (int p = 0; p < N; p++) dst[d->perm[p]] = src[p];
    };
    if (grad_friction) {
        HIPCHK(d, hipMemcpy(h.data(), a->nbar, N * sizeof(double), hipMemcpyDeviceToHost));
        unpermute(h.data(), grad_friction);
    }
    if (grad_elevation) {
        HIPCHK(d, hipMemcpy(h.data(), a->zbar, N * sizeof(double), hipMemcpyDeviceToHost));
        unpermute(h.data(), grad_elevation);
    }
    if (grad_state0) {
        HIPCHK(d, hipMemcpy(h.data(), a->lam, n3 * sizeof(double), hipMemcpyDeviceToHost));
        for (int q = 0; q < 3; q++) unpermute(h.data() + (size_t)q * N, grad_state0 + (size_t)q * N);
    }
    return ANUGA_OK;
}

}  // extern "C"

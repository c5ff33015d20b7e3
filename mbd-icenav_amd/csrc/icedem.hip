// icedem.hip -- kernels and C ABI of the MI355X DPDEM ice-ship frame step.
// Build: see Makefile (hipcc --offload-arch=gfx950 -O3 -ffp-contract=off).
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>
#include <algorithm>
#include <cmath>

#include "frame_kernels.h"

using namespace icedem;

// ═════════════════════════════════════════════════════════════════════════════
// Kernels
// ═════════════════════════════════════════════════════════════════════════════

// ── leaf kernels ─────────────────────────────────────────────────────────────

__global__ __launch_bounds__(BLOCK) void k_collisions_leaf(
    const float2* __restrict__ polys, const uint8_t* __restrict__ mask, const float* __restrict__ pos,
    const float* __restrict__ rot, const float* __restrict__ bound, int n, int maxv,
    float sx, float sy, float sh, const float2* __restrict__ hull, int S, IceDemParams prm,
    float* pen, float* nx, float* ny, float* cx, float* cy) {
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    const int wave_base = blockIdx.x * BLOCK + (threadIdx.x & ~63);
    float px = 0.0f, py = 0.0f, ro = 0.0f;
    if (e < n) { px = pos[e]; py = pos[n + e]; ro = rot[e]; }
    float a, b, c, d, f;
    hull_wave(prm, polys, mask, bound, hull, S, maxv, wave_base, n, px, py, ro, sx, sy, sh, a, b, c, d, f, true);
    if (e < n) { pen[e] = a; nx[e] = b; ny[e] = c; cx[e] = d; cy[e] = f; }
}

__global__ void k_bound_radius(const float2* __restrict__ polys, int n, int maxv, float* bound) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    float m = 0.0f;
    for (int v = 0; v < maxv; ++v) {
        float2 p = polys[(size_t)e * maxv + v];
        m = fmaxf(m, sqrtf(p.x * p.x + p.y * p.y));
    }
    bound[e] = m * 1.0001f + 1e-3f;
}

__global__ void k_bond_damage_leaf(int nb, const float* dn, const float* ds, const float* sn, const float* st,
                                   const float* sb, float C, float mu, float g1, float g2, float eta,
                                   const float* dmg, const float* dmx, const float* dm0, const uint8_t* intact,
                                   float* o_d, float* o_mx, float* o_m0, uint8_t* o_int) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb) return;
    float a, b, c;
    bool d;
    bond_damage_update(dn[t], ds[t], sn[t], st[t], sb[t], C, mu, g1, g2, eta, dmg[t], dmx[t], dm0[t],
                       intact[t] != 0, a, b, c, d);
    o_d[t] = a; o_mx[t] = b; o_m0[t] = c; o_int[t] = d ? 1 : 0;
}

__global__ void k_cc_copy(int n, const int32_t* a, int32_t* b) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = a[t];
}

// one sweep of find_connected_components_jax (jax_bonds.py:72-86): reads old
// labels, scatter-min into new (integer min: order-free)
__global__ void k_cc_sweep(int n, int nb, const int32_t* bi, const int32_t* bj, const uint8_t* intact,
                           const int32_t* old_l, int32_t* new_l) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb) return;
    int li = old_l[bi[t]], lj = old_l[bj[t]];
    int m = min(li, lj);
    if (!intact[t]) m = n;
    atomicMin(&new_l[bi[t]], m);
    atomicMin(&new_l[bj[t]], m);
}

// labels[i] = i: the starting labels of find_connected_components_jax
// (jax_bonds.py:68)
__global__ void k_cc_init(int n, int32_t* labels) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) labels[t] = t;
}

__global__ void k_iota_count(int n, const int32_t* labels, int32_t* count) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) atomicAdd(&count[labels[t]], 1);
}

// Stable placement of elements grouped by label (one workgroup, chunks of
// 1024 in element order) so per-label sums run in element order.
__global__ __launch_bounds__(1024) void k_stable_place(int n, const int32_t* labels, const int32_t* start,
                                                       int32_t* fill, int32_t* order) {
    __shared__ int lab[1024];
    for (int base = 0; base < n; base += 1024) {
        const int t = base + threadIdx.x;
        const int L = t < n ? labels[t] : -1;
        lab[threadIdx.x] = L;
        __syncthreads();
        int rank = 0;
        bool last = true;
        if (t < n) {
            const int lim = min(1024, n - base);
            for (int q = 0; q < lim; ++q) {
                if (lab[q] == L) {
                    if (q < (int)threadIdx.x) ++rank;
                    else if (q > (int)threadIdx.x) last = false;
                }
            }
            order[start[L] + fill[L] + rank] = t;
        }
        __syncthreads();
        if (t < n && last) fill[L] += rank + 1;
        __syncthreads();
    }
}

__global__ void k_com_sum(int n, const int32_t* start, const int32_t* order, const float* masses,
                          const float* vel, float* tm, float* tx, float* ty) {
    const int L = blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= n) return;
    float m = 0.0f, x = 0.0f, y = 0.0f;
    for (int q = start[L]; q < start[L + 1]; ++q) {
        const int e = order[q];
        const float me = masses[e];
        m = m + me;
        x = x + me * vel[e];
        y = y + me * vel[n + e];
    }
    tm[L] = m; tx[L] = x; ty[L] = y;
}

__global__ void k_com_out(int n, const int32_t* labels, const float* tm, const float* tx, const float* ty,
                          float* out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int L = labels[e];
    const float sm = vmax(tm[L], 1e-30f);
    out[e] = tx[L] / sm;
    out[n + e] = ty[L] / sm;
}

__global__ void k_excl_scan_small(int m, const int32_t* in, int32_t* out) {
    // one thread: exclusive scan of m+1 entries (leaf use only)
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int run = 0;
    for (int i = 0; i < m; ++i) { out[i] = run; run += in[i]; }
    out[m] = run;
}

__global__ __launch_bounds__(1024) void k_excl_scan(int m, const int32_t* in, int32_t* out) {
    __shared__ int lds[1024];
    const int T = blockDim.x, tid = threadIdx.x;
    const int chunk = (m + T - 1) / T;
    const int lo = min(m, tid * chunk), hi = min(m, lo + chunk);
    int s = 0;
    for (int i = lo; i < hi; ++i) s += in[i];
    lds[tid] = s;
    __syncthreads();
    for (int off = 1; off < T; off <<= 1) {
        int v = (tid >= off) ? lds[tid - off] : 0;
        __syncthreads();
        lds[tid] += v;
        __syncthreads();
    }
    int run = lds[tid] - s;
    for (int i = lo; i < hi; ++i) { out[i] = run; run += in[i]; }
    if (tid == T - 1) out[m] = lds[T - 1];
}

// Static per-entry bond parameters of the bond table: rest length and the
// stiffness E * (h_i + h_j) / 2 * factor of jax_simulation.py:63-108, with
// the reference's operation order.
__global__ void k_bond_prm(const DevView* __restrict__ V) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int W = V->W;
    if (W <= 0 || idx >= (size_t)V->n * W) return;
    const int e = (int)(idx / W), q = (int)(idx % W);
    float2 out = make_float2(0.0f, 0.0f);
    if (q < V->bdeg[e]) {
        const int b = V->btab[idx].x;
        const int i = V->bi[b], j = V->bj[b];
        const float avg_h = (V->thick[i] + V->thick[j]) * 0.5f;
        const float stiff = V->p.youngs_modulus * avg_h * V->p.bond_stiffness_factor;
        out = make_float2(V->rest[b], stiff);
    }
    const_cast<float2*>(V->bprm)[idx] = out;
}

// bond-major stiffness (the same float operations as k_bond_prm)
__global__ void k_bond_stiff(const DevView* __restrict__ V) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= V->nb) return;
    const int i = V->bi[b], j = V->bj[b];
    const float avg_h = (V->thick[i] + V->thick[j]) * 0.5f;
    const_cast<float*>(V->bstf)[b] = V->p.youngs_modulus * avg_h * V->p.bond_stiffness_factor;
}

// Contact pairs (e < j) of element e in ascending j, by repeated selection
// over the scanned cells of both grids and the overflow list.
__device__ int pairs_of(const DevView* __restrict__ V, const float* __restrict__ pos, int e,
                        int32_t* pi, int32_t* pj, int64_t o, int64_t max_pairs) {
    const IceDemParams& P = V->p;
    const int n = V->n, cap = P.grid_bucket;
    int x0, x1, y0, y1, u0 = 0, u1 = -1, v0 = 0, v1 = -1;
    scan_range(P, V->radius[e], pos[e], pos[n + e], x0, x1, y0, y1);
    if (V->nbig > 0) scan_range2(V, pos[e], pos[n + e], u0, u1, v0, v1);
    const int novf = V->ovf_count[0];
    int last = e, c = 0;
    for (;;) {
        int best = 0x7fffffff;
        auto consider = [&](int j) {
            if (j > last && j < best && pair_overlap(P, G(pos), G(V->radius), n, e, j)) best = j;
        };
        for (int gy = y0; gy <= y1; ++gy)
            for (int gx = x0; gx <= x1; ++gx) {
                const int cell = gy * P.grid_nx + gx;
                GAS const uint8_t* rec = rec_ptr(V, fine_rec(V, 0, cell));
                const int cc = min(*(GAS const int*)rec, cap);
                for (int s = 0; s < cc; ++s) consider(slot_load(rec, s).e);
            }
        for (int gy = v0; gy <= v1; ++gy)
            for (int gx = u0; gx <= u1; ++gx) {
                const int cell = gy * V->nx2 + gx;
                GAS const uint8_t* rec = rec_ptr(V, coarse_rec(V, 0, cell));
                const int cc = min(*(GAS const int*)rec, V->cap2);
                for (int s = 0; s < cc; ++s) consider(slot_load(rec, s).e);
            }
        for (int s = 0; s < novf; ++s) consider(V->ovf_list[s]);
        if (best == 0x7fffffff) break;
        if (pi && o + c < max_pairs) { pi[o + c] = e; pj[o + c] = best; }
        last = best;
        ++c;
    }
    return c;
}

__global__ __launch_bounds__(BLOCK) void k_pair_count(const DevView* __restrict__ V, const float* __restrict__ pos,
                                                      int32_t* cnt) {
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= V->n) return;
    cnt[e] = pairs_of(V, pos, e, nullptr, nullptr, 0, 0);
}

__global__ __launch_bounds__(BLOCK) void k_pair_emit(const DevView* __restrict__ V, const float* __restrict__ pos,
                                                     const int32_t* off, int64_t max_pairs, int32_t* pi, int32_t* pj) {
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= V->n) return;
    pairs_of(V, pos, e, pi, pj, off[e], max_pairs);
}

__global__ void k_ship_leaf(const DevView* __restrict__ V, float* ship, const float* sp, const float* tau,
                            const float* vws) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float spc[3] = {sp[0], sp[1], sp[2]}, spn[3] = {sp[3], sp[4], sp[5]};
    float t3[3] = {tau[0], tau[1], tau[2]};
    float st[ICEDEM_SHIP_STATE_LEN];
    for (int i = 0; i < ICEDEM_SHIP_STATE_LEN; ++i) st[i] = ship[i];
    ship_step(V->p, st, spc, spn, t3, vws ? vws[0] : 0.0f, vws ? vws[1] : 0.0f);
    for (int i = 0; i < ICEDEM_SHIP_STATE_LEN; ++i) ship[i] = st[i];
}

// Ship-only rollouts scored on a costmap (the costmap planners' rollout,
// jax_mbd_planner.py:569-660): ship_mode 1 dynamics with no ice forces; at
// every waypoint (start and after each frame) the costmap is read bilinearly
// (jax_mbd_planner.py:128-168) at the ship's centre and four hull corners and
// the maximum kept; segments add 0.5 (c_t + c_t+1) |d| with |d| =
// sqrt(dx^2 + dy^2 + 1e-8).  One thread per candidate.
__device__ __forceinline__ float cmap_query(const float* __restrict__ cm, int rows, int cols, float x0, float y0,
                                            float res, float oob, float x, float y) {
    const float cf = (x - x0) / res - 0.5f, rf = (y - y0) / res - 0.5f;
    const float r0f = floorf(rf), c0f = floorf(cf);
    const int r0 = (int)fminf(fmaxf(r0f, -2.0f), (float)rows), c0 = (int)fminf(fmaxf(c0f, -2.0f), (float)cols);
    const float fr = rf - r0f, fc = cf - c0f;
    const int r0c = min(max(r0, 0), rows - 1), r1c = min(max(r0 + 1, 0), rows - 1);
    const int c0c = min(max(c0, 0), cols - 1), c1c = min(max(c0 + 1, 0), cols - 1);
    const float pen = (c0 < 0 || c0 + 1 >= cols) ? oob : 0.0f;
    const float v = cm[(size_t)r0c * cols + c0c] * (1.0f - fr) * (1.0f - fc) + cm[(size_t)r0c * cols + c1c] * (1.0f - fr) * fc +
                    cm[(size_t)r1c * cols + c0c] * fr * (1.0f - fc) + cm[(size_t)r1c * cols + c1c] * fr * fc;
    return v + pen;
}

__device__ __forceinline__ float hull_cost(const float* __restrict__ cm, int rows, int cols, float x0, float y0,
                                           float res, float oob, float hl, float hw, const float* st) {
    float s, c;
    sincosf(st[2], &s, &c);
    const float bx[5] = {0.0f, hl, hl, -hl, -hl}, by[5] = {0.0f, hw, -hw, hw, -hw};
    float m = -INFINITY;
    #pragma unroll
    for (int i = 0; i < 5; ++i)
        m = fmaxf(m, cmap_query(cm, rows, cols, x0, y0, res, oob, st[0] + c * bx[i] - s * by[i],
                                st[1] + s * bx[i] + c * by[i]));
    return m;
}

__global__ void k_ship_rollouts(const DevView* __restrict__ V, const float* __restrict__ ship0, int ship0_stride,
                                const float* __restrict__ ctl, int S, int nk, int hold, int frames,
                                const float* __restrict__ cm, int rows, int cols, float x0, float y0, float res,
                                float oob, float hl, float hw, float* __restrict__ out, int snap_every, int n_snap,
                                float* __restrict__ snap) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= S) return;
    const IceDemParams& P = V->p;
    float st[ICEDEM_SHIP_STATE_LEN];
    for (int i = 0; i < ICEDEM_SHIP_STATE_LEN; ++i) st[i] = ship0[(size_t)r * ship0_stride + i];
    const float ke0 = 0.5f * P.ship_mass * (st[3] * st[3] + st[4] * st[4]) + 0.5f * P.ship_Izz * (st[5] * st[5]);
    const float zero3[3] = {0.0f, 0.0f, 0.0f};
    float wp = cm ? hull_cost(cm, rows, cols, x0, y0, res, oob, hl, hw, st) : 0.0f;
    float cost = 0.0f, len = 0.0f;
    const float* c = ctl + (size_t)r * nk * 2;
    const float sx0 = st[0], sy0 = st[1], psi0 = st[2];
    float s0, c0;
    sincosf(psi0, &s0, &c0);
    for (int k = 0; k < frames; ++k) {
        const int knot = min(k / max(hold, 1), nk - 1);
        const float px = st[0], py = st[1];
        ship_step(P, st, zero3, zero3, zero3, 0.0f, 0.0f, c[2 * knot], c[2 * knot + 1]);
        const float dx = st[0] - px, dy = st[1] - py;
        const float seg = sqrtf(dx * dx + dy * dy + 1e-8f);
        const float w2 = cm ? hull_cost(cm, rows, cols, x0, y0, res, oob, hl, hw, st) : 0.0f;
        cost = cost + 0.5f * (wp + w2) * seg;
        len = len + seg;
        wp = w2;
        if (snap && snap_every > 0 && (k + 1) % snap_every == 0 && (k + 1) / snap_every <= n_snap) {
            // position in the start's ship frame, heading change, speed
            float* o = snap + ((size_t)r * n_snap + ((k + 1) / snap_every - 1)) * 4;
            const float ex = st[0] - sx0, ey = st[1] - sy0;
            o[0] = c0 * ex + s0 * ey;
            o[1] = -s0 * ex + c0 * ey;
            o[2] = st[2] - psi0;
            o[3] = sqrtf(st[3] * st[3] + st[4] * st[4]);
        }
    }
    const float ke = 0.5f * P.ship_mass * (st[3] * st[3] + st[4] * st[4]) + 0.5f * P.ship_Izz * (st[5] * st[5]);
    out[4 * r] = cost;
    out[4 * r + 1] = ke0 - ke;
    out[4 * r + 2] = len;
    out[4 * r + 3] = st[2];
}

// ═════════════════════════════════════════════════════════════════════════════
// Host side
// ═════════════════════════════════════════════════════════════════════════════

// The scene's static tables (polygons, masses, bonds, bond table and
// parameters, hull): read-only after creation, shared by every context made
// with icedem_create_like and freed with the last of them.
struct StaticTables {
    int device = 0;
    std::vector<void*> allocs;
    ~StaticTables() {
        (void)hipSetDevice(device);
        for (void* p : allocs) (void)hipFree(p);
    }
};

struct IceDemCtx {
    int device = 0;
    DevView hv{};              // host copy of the device view
    DevView* dv = nullptr;     // device view
    CallDesc* dd = nullptr;    // device call descriptor
    CallDesc* hd = nullptr;    // pinned host staging for the descriptor
    hipEvent_t desc_done = nullptr;
    std::shared_ptr<StaticTables> statics;
    std::vector<void*> allocs; // this context's own buffers
    std::map<int, hipGraphExec_t> graphs;
    std::map<int, hipGraph_t> graph_defs;
    std::string err;
};

static thread_local std::string g_err;

#define HIP_TRY(expr)                                                                         \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) {                                                               \
            set_err(ctx, std::string(#expr) + ": " + hipGetErrorString(_e));                  \
            return ICEDEM_E_HIP;                                                              \
        }                                                                                     \
    } while (0)

static void set_err(IceDemCtx* ctx, const std::string& s) {
    if (ctx) ctx->err = s;
    g_err = s;
}

// error text for the surrogate entry points (surrogate.hip)
void icedem_set_error(const std::string& msg) { set_err(nullptr, msg); }

// device buffer of count T (at least one) owned by the context, or with
// stat by its static tables
template <class T>
static int dalloc(IceDemCtx* ctx, T** p, size_t count, bool stat = false) {
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    void* q = nullptr;
    HIP_TRY(hipMalloc(&q, bytes));
    (stat ? ctx->statics->allocs : ctx->allocs).push_back(q);
    *p = reinterpret_cast<T*>(q);
    return 0;
}

template <class T>
static int dupload(IceDemCtx* ctx, T** p, const T* src, size_t count, bool stat = false) {
    int rc = dalloc(ctx, p, count, stat);
    if (rc) return rc;
    if (count) HIP_TRY(hipMemcpy(*p, src, count * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

static int alloc_state_buffers(IceDemCtx* ctx);

extern "C" {

const char* icedem_status_string(int code) {
    switch (code) {
        case ICEDEM_OK: return "ok";
        case ICEDEM_E_INVALID: return "invalid argument";
        case ICEDEM_E_HIP: return "HIP runtime error";
        case ICEDEM_E_LIMIT: return "size beyond a documented limit";
        case ICEDEM_E_DEVICE: return "error flagged by a kernel";
        default: return "unknown status";
    }
}

const char* icedem_last_error(const IceDemCtx* ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int icedem_version(void) { return 1; }

int icedem_destroy(IceDemCtx* ctx) {
    if (!ctx) return ICEDEM_OK;
    (void)hipSetDevice(ctx->device);
    for (auto& g : ctx->graphs) (void)hipGraphExecDestroy(g.second);
    for (auto& g : ctx->graph_defs) (void)hipGraphDestroy(g.second);
    for (void* p : ctx->allocs) (void)hipFree(p);
    if (ctx->hd) (void)hipHostFree(ctx->hd);
    if (ctx->desc_done) (void)hipEventDestroy(ctx->desc_done);
    delete ctx;
    return ICEDEM_OK;
}

int icedem_create(const IceDemStatic* st, const IceDemParams* prm, int device, IceDemCtx** out) {
    IceDemCtx* ctx = nullptr;
    if (!st || !prm || !out) { set_err(nullptr, "null argument"); return ICEDEM_E_INVALID; }
    *out = nullptr;
    if (st->n < 1 || st->n_bonds < 0 || st->n_floes < 0 || st->max_v < 1 || st->n_hull < 1) {
        set_err(nullptr, "bad sizes"); return ICEDEM_E_INVALID;
    }
    if (st->max_v > ICEDEM_MAX_VERTS) { set_err(nullptr, "max_v > 64"); return ICEDEM_E_LIMIT; }
    if (prm->n_edge_samples < 0 || prm->n_edge_samples > ICEDEM_MAX_EDGE_SAMPLES) {
        set_err(nullptr, "n_edge_samples out of range"); return ICEDEM_E_LIMIT;
    }
    const int n = st->n;
    const int nblk = (n + BLOCK - 1) / BLOCK;
    int nbpad = 1;
    while (nbpad < nblk) nbpad <<= 1;
    if (nbpad > MAX_PARTIALS) { set_err(nullptr, "n > 1048576 elements"); return ICEDEM_E_LIMIT; }
    if (prm->grid_nx < 1 || prm->grid_ny < 1 || (int64_t)prm->grid_nx * prm->grid_ny > (1 << 26) ||
        !(prm->grid_cell > 0.0f)) {
        set_err(nullptr, "bad grid"); return ICEDEM_E_INVALID;
    }
    if (prm->n_ice_steps < 1) { set_err(nullptr, "n_ice_steps < 1"); return ICEDEM_E_INVALID; }
    if (prm->force_lanes != 0 && prm->force_lanes != 8 && prm->force_lanes != 16) {
        set_err(nullptr, "force_lanes must be 0, 8 or 16"); return ICEDEM_E_INVALID;
    }
    for (int b = 0; b < st->n_bonds; ++b) {
        if (st->h_bond_i[b] < 0 || st->h_bond_i[b] >= n || st->h_bond_j[b] < 0 || st->h_bond_j[b] >= n) {
            set_err(nullptr, "bond index out of range"); return ICEDEM_E_INVALID;
        }
    }
    for (int e = 0; e < n; ++e) {
        if (st->h_floe_of[e] < 0 || st->h_floe_of[e] >= st->n_floes) {
            set_err(nullptr, "floe index out of range"); return ICEDEM_E_INVALID;
        }
    }
    ctx = new IceDemCtx();
    ctx->device = device;
    ctx->statics = std::make_shared<StaticTables>();
    ctx->statics->device = device;
    HIP_TRY(hipSetDevice(device));
    DevView& H = ctx->hv;
    H.p = *prm;
    H.n = n; H.nb = st->n_bonds; H.nf = st->n_floes; H.maxv = st->max_v; H.S = st->n_hull;
    H.G = prm->grid_nx * prm->grid_ny;
    H.nblk = nblk; H.nbpad = nbpad;
    H.lanes = resolve_lanes(prm->force_lanes, n);
    if (!(prm->list_skin >= 0.0f) || !std::isfinite(prm->list_skin)) {
        set_err(nullptr, "list_skin must be finite and >= 0"); delete ctx; return ICEDEM_E_INVALID;
    }
    {   // partner lists: rebuild once 2 x (largest displacement) may reach the
        // skin less a margin far above the float32 rounding of distances
        const float thr = 0.5f * (prm->list_skin - 0.01f);
        H.nl_on = (prm->n_ice_steps == 1 && thr > 0.0f) ? 1 : 0;
        H.nl_skin = H.nl_on ? prm->list_skin : 0.0f;
        H.nl_thr2 = H.nl_on ? thr * thr : 0.0f;
    }
    H.bf_on = (prm->n_ice_steps == 1 && st->n_bonds > 0) ? 1 : 0;
    {   // near candidates: the near test's 1 m margin plus one frame's largest
        // element displacement (velocity clipped per axis), the hull box's
        // shift at the clipped ship speed and the turn of its far corner at
        // the clipped yaw rate, and 0.15 m (the ship's unclipped velocity of a
        // step); a near element outside the candidates is still tested
        double maxb = 0.0;
        for (int e = 0; e < n; ++e) maxb = std::max(maxb, (double)st->h_bound_radius[e]);
        const double R = std::hypot((double)prm->ship_box_half_len + maxb + 1.0,
                                    (double)prm->ship_box_half_beam + maxb + 1.0);
        H.cand_margin = (float)(1.0 + (double)prm->dt * ((double)prm->velocity_clip * std::sqrt(2.0) +
                                                         std::hypot((double)prm->vc_surge, (double)prm->vc_sway) +
                                                         (double)prm->vc_yaw * R) + 0.15);
        if (prm->near_cand_margin >= 0.0f) H.cand_margin = prm->near_cand_margin;
    }
    if (prm->grid_bucket < 1 || (int64_t)H.G * prm->grid_bucket > (1LL << 28)) {
        set_err(nullptr, "bad grid_bucket"); delete ctx; return ICEDEM_E_INVALID;
    }
    {   // coarse grid for the elements above grid_rsmall: cell2 >= 2 r_max + delta
        double rmax = 0.0;
        int nbig = 0;
        for (int e = 0; e < n; ++e) {
            const float r = st->h_radii[e];
            if (!(r >= 0.0f) || !std::isfinite(r)) { set_err(nullptr, "bad radius"); delete ctx; return ICEDEM_E_INVALID; }
            rmax = std::max(rmax, (double)r);
            if (r > prm->grid_rsmall) ++nbig;
        }
        // contacts reach at most 2 r_max + delta; with cell2 >= 2 (2 r_max + delta)
        // the 2x2 coarse cells on an element's side hold all its large partners
        // (list builds reach list_skin further)
        const double c2 = std::max((double)prm->grid_cell,
                                   2.0 * (2.0 * rmax + prm->contact_dilation2 + H.nl_skin) * 1.002 + 1e-3);
        H.dreach = (float)((2.0 * rmax + prm->contact_dilation2 + H.nl_skin) * 1.001 + 1e-3);
        H.cell2 = (float)c2;
        if ((double)H.cell2 < c2) H.cell2 = std::nextafter(H.cell2, INFINITY);
        H.icell2 = 1.0f / H.cell2;
        H.nx2 = std::max(1, (int)std::ceil((double)prm->grid_nx * prm->grid_cell / H.cell2));
        H.ny2 = std::max(1, (int)std::ceil((double)prm->grid_ny * prm->grid_cell / H.cell2));
        H.G2 = H.nx2 * H.ny2;
        H.nbig = nbig;
        H.cap2 = prm->grid_bucket2 > 0 ? std::min(prm->grid_bucket2, std::max(nbig, 1)) : std::max(1, std::min(nbig, 8));
        if ((int64_t)H.G2 * H.cap2 > (1LL << 28)) { set_err(nullptr, "bad grid_bucket2"); delete ctx; return ICEDEM_E_INVALID; }
        H.ovf_cap = n + 64;   // every element is inserted once
    }
    const size_t nv = (size_t)n * st->max_v;
    int rc = 0;
#define UP(field, type, src, cnt) do { type* _p; rc = dupload(ctx, &_p, (const type*)(src), (cnt), true); if (rc) { icedem_destroy(ctx); return rc; } H.field = _p; } while (0)
    UP(polys, float2, st->h_local_polys, nv);
    UP(mask, uint8_t, st->h_poly_mask, nv);
    UP(mass, float, st->h_masses, n);
    UP(inertia, float, st->h_inertia, n);
    UP(radius, float, st->h_radii, n);
    UP(area, float, st->h_areas, n);
    UP(thick, float, st->h_thicknesses, n);
    UP(bound, float, st->h_bound_radius, n);
    UP(floe, int32_t, st->h_floe_of, n);
    UP(bi, int32_t, st->h_bond_i, st->n_bonds);
    UP(bj, int32_t, st->h_bond_j, st->n_bonds);
    {
        std::vector<int2> bp(std::max(1, st->n_bonds));
        for (int b = 0; b < st->n_bonds; ++b) {
            const int i = st->h_bond_i[b], j = st->h_bond_j[b];
            const bool brk = (st->h_thicknesses[i] <= prm->max_fracture_thickness) &&
                             (st->h_thicknesses[j] <= prm->max_fracture_thickness);
            bp[b] = make_int2(st->h_floe_of[i], st->h_floe_of[j] | (brk ? 0 : (1 << 30)));
        }
        UP(bfloe, int2, bp.data(), bp.size());
    }
    UP(rest, float, st->h_rest_lengths, st->n_bonds);
    if (st->n_bonds > 0) {   // bond-major stiffness, the formula of k_bond_prm
        float* bs;
        rc = dalloc(ctx, &bs, st->n_bonds, true);
        if (rc) { icedem_destroy(ctx); return rc; }
        H.bstf = bs;
    }
    UP(bstr, float, st->h_bond_strengths, st->n_bonds);
    UP(sbar, float, st->h_sigma_bar, st->n_bonds);
    UP(hull, float2, st->h_hull_local, st->n_hull);
    // element-major bond table in the reference's scatter order: bonds where
    // the element is i (ascending), then bonds where it is j (ascending)
    {
        std::vector<int32_t> deg(n, 0);
        for (int b = 0; b < st->n_bonds; ++b) { deg[st->h_bond_i[b]]++; deg[st->h_bond_j[b]]++; }
        int W = 0;
        for (int e = 0; e < n; ++e) W = std::max(W, deg[e]);
        if ((int64_t)n * W > (1LL << 30)) {
            set_err(nullptr, "bond table beyond 2^30 entries"); icedem_destroy(ctx); return ICEDEM_E_LIMIT;
        }
        H.W = W;
        const size_t tw = std::max<size_t>(1, (size_t)n * W);
        std::vector<int4> tab(tw, make_int4(0, 0, 0, 0));
        std::vector<int32_t> fill(n, 0);
        for (int b = 0; b < st->n_bonds; ++b) {
            const int i = st->h_bond_i[b], j = st->h_bond_j[b];
            tab[(size_t)i * W + fill[i]++] = make_int4(b, j, st->h_floe_of[j], 0);
        }
        for (int b = 0; b < st->n_bonds; ++b) {
            const int i = st->h_bond_i[b], j = st->h_bond_j[b];
            tab[(size_t)j * W + fill[j]++] = make_int4(b, i, st->h_floe_of[i], 1);
        }
        UP(bdeg, int32_t, deg.data(), n);
        UP(btab, int4, tab.data(), tw);
        float2* bp;
        rc = dalloc(ctx, &bp, tw, true);
        if (rc) { icedem_destroy(ctx); return rc; }
        H.bprm = bp;
    }
#undef UP
    {   // bucket record sizes
        auto rec_bytes = [](int64_t cap) { return ((16 + 24 * cap + 127) / 128) * 128; };
        H.rec_f = rec_bytes(prm->grid_bucket);
        H.rec_c = rec_bytes(H.cap2);
        H.coarse_off = 2 * (int64_t)H.G * H.rec_f;
        const int64_t total = H.coarse_off + 2 * (int64_t)H.G2 * H.rec_c;
        if (2 * (int64_t)H.G + 2 * (int64_t)H.G2 > (int64_t)INT32_MAX || total >= (1LL << 32)) {
            set_err(nullptr, "bucket records beyond the addressable range"); icedem_destroy(ctx); return ICEDEM_E_LIMIT;
        }
    }
    rc = alloc_state_buffers(ctx);
    if (rc) { icedem_destroy(ctx); return rc; }
    if ((size_t)n * H.W > 0) {
        const size_t tw = (size_t)n * H.W;
        hipLaunchKernelGGL(k_bond_prm, dim3((unsigned)((tw + 255) / 256)), dim3(256), 0, 0, (const DevView*)ctx->dv);
        HIP_TRY(hipGetLastError());
    }
    if (st->n_bonds > 0) {
        hipLaunchKernelGGL(k_bond_stiff, dim3((unsigned)((st->n_bonds + 255) / 256)), dim3(256), 0, 0,
                           (const DevView*)ctx->dv);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipDeviceSynchronize());
    *out = ctx;
    return ICEDEM_OK;
}

int icedem_create_like(IceDemCtx* base, IceDemCtx** out) {
    if (!base || !out) { set_err(nullptr, "null argument"); return ICEDEM_E_INVALID; }
    *out = nullptr;
    IceDemCtx* ctx = new IceDemCtx();
    ctx->device = base->device;
    ctx->statics = base->statics;
    ctx->hv = base->hv;                 // sizes, parameters and the static tables' pointers
    HIP_TRY(hipSetDevice(ctx->device));
    const int rc = alloc_state_buffers(ctx);
    if (rc) { icedem_destroy(ctx); return rc; }
    HIP_TRY(hipDeviceSynchronize());
    *out = ctx;
    return ICEDEM_OK;
}

}  // extern "C"

// The context's own buffers (per-frame scratch, buckets, counters, next
// state), zeroed where the frame kernels expect zeros, and the device view.
static int alloc_state_buffers(IceDemCtx* ctx) {
    DevView& H = ctx->hv;
    const int n = H.n, nblk = H.nblk, nbpad = H.nbpad;
    int rc = 0;
#define AL(field, type, cnt) do { type* _p; rc = dalloc(ctx, &_p, (cnt)); if (rc) return rc; H.field = _p; } while (0)
    AL(pen, float, n); AL(nx, float, n); AL(ny, float, n); AL(fm_raw, float, n);
    AL(soi_x, float, n); AL(soi_y, float, n); AL(mterm, float, n); AL(fm, float, n);
    AL(shipf, float, 2 * (size_t)n);
    AL(nearf, uint8_t, n); AL(near_list, int32_t, n);
    AL(cand, int32_t, 2 * (size_t)n); AL(cflag, uint8_t, 2 * (size_t)n);
    AL(nearcnt, int32_t, nbpad); AL(counters, int32_t, 4);
    AL(crack, uint8_t, n);
    AL(cells, uint8_t, (size_t)(H.coarse_off + 2 * (int64_t)H.G2 * H.rec_c));   // fine sets, then coarse
    AL(ovf_list, int32_t, 2 * (size_t)H.ovf_cap); AL(ovf_count, int32_t, 2);
    AL(floe_coll, uint32_t, 2 * (size_t)H.nf); AL(floe_maxfm, uint32_t, 2 * (size_t)H.nf);
    AL(part1, float, 3 * (size_t)nbpad); AL(part2, float, 2 * (size_t)nbpad); AL(scal, float, 16); AL(shpre, float, 32);
    AL(nxt_pos, float, 2 * (size_t)n); AL(nxt_vel, float, 2 * (size_t)n);
    AL(nxt_rot, float, n); AL(nxt_omega, float, n);
    AL(err, int32_t, 4);
    if (H.bf_on) AL(bfv, float4, (size_t)H.nb);
    if (H.nl_on) {
        AL(nl, int32_t, (size_t)n * NLW); AL(nl_cnt, int32_t, n); AL(xb, float, 2 * (size_t)n);
        AL(es, float, 6 * (size_t)n); AL(nlctl, uint32_t, NLCTL);
        AL(dstamp, uint32_t, (size_t)H.G2); AL(rneed, uint8_t, n);
    }
#undef AL
    HIP_TRY(hipMemset(H.cells, 0, (size_t)(H.coarse_off + 2 * (int64_t)H.G2 * H.rec_c)));
    HIP_TRY(hipMemset(H.ovf_count, 0, sizeof(int32_t) * 2));
    HIP_TRY(hipMemset(H.counters, 0, sizeof(int32_t) * 4));
    HIP_TRY(hipMemset(H.nearcnt, 0, sizeof(int32_t) * nbpad));   // [nblk, nbpad) stays 0
    HIP_TRY(hipMemset(H.cflag, 0, 2 * (size_t)n));
    HIP_TRY(hipMemset(H.floe_coll, 0, sizeof(uint32_t) * 2 * std::max(H.nf, 1)));
    HIP_TRY(hipMemset(H.floe_maxfm, 0, sizeof(uint32_t) * 2 * std::max(H.nf, 1)));
    HIP_TRY(hipMemset(H.part1, 0, sizeof(float) * 3 * nbpad));
    HIP_TRY(hipMemset(H.part2, 0, sizeof(float) * 2 * nbpad));
    HIP_TRY(hipMemset(H.scal, 0, sizeof(float) * 16));
    HIP_TRY(hipMemset(H.err, 0, sizeof(int32_t) * 4));
    if (H.nl_on) {
        HIP_TRY(hipMemset(H.nlctl, 0, sizeof(uint32_t) * NLCTL));
        HIP_TRY(hipMemset(H.es, 0, sizeof(float) * 6 * (size_t)n));
        HIP_TRY(hipMemset(H.dstamp, 0, sizeof(uint32_t) * (size_t)H.G2));
        HIP_TRY(hipMemset(H.rneed, 1, n));
        // the first call builds: call-start slot at +inf
        const uint32_t inf[NLSUB] = {0x7f800000u, 0x7f800000u, 0x7f800000u, 0x7f800000u,
                                     0x7f800000u, 0x7f800000u, 0x7f800000u, 0x7f800000u};
        HIP_TRY(hipMemcpy(H.nlctl + 4 * NLSUB, inf, sizeof(inf), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemset(H.nearf, 0, n));
    HIP_TRY(hipMemset(H.nearcnt, 0, sizeof(int32_t) * nblk));
    {
        DevView* d;
        rc = dalloc(ctx, &d, 1);
        if (rc) return rc;
        ctx->dv = d;
        CallDesc* cd;
        rc = dalloc(ctx, &cd, 1);
        if (rc) return rc;
        ctx->dd = cd;
    }
    HIP_TRY(hipMemcpy(ctx->dv, &H, sizeof(DevView), hipMemcpyHostToDevice));
    HIP_TRY(hipHostMalloc((void**)&ctx->hd, sizeof(CallDesc), hipHostMallocDefault));
    memset(ctx->hd, 0, sizeof(CallDesc));
    HIP_TRY(hipEventCreateWithFlags(&ctx->desc_done, hipEventDisableTiming));
    return ICEDEM_OK;
}

extern "C" {

static int validate_state(IceDemCtx* ctx, const IceDemState* s) {
    if (!s || !s->d_pos || !s->d_vel || !s->d_rot || !s->d_omega || !s->d_sub || !s->d_active || !s->d_ship ||
        (ctx->hv.nf > 0 && !s->d_floe_elig) ||
        (ctx->hv.nb > 0 && (!s->d_bond_intact || !s->d_bond_damage || !s->d_bond_dm0 || !s->d_bond_dmmax ||
                            !s->d_bond_ds))) {
        set_err(ctx, "state has a null pointer");
        return ICEDEM_E_INVALID;
    }
    return ICEDEM_OK;
}

static int put_desc(IceDemCtx* ctx, const CallDesc& cd, hipStream_t stream) {
    // the previous descriptor copy must have left the pinned buffer
    HIP_TRY(hipEventSynchronize(ctx->desc_done));
    *ctx->hd = cd;
    HIP_TRY(hipMemcpyAsync(ctx->dd, ctx->hd, sizeof(CallDesc), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipEventRecord(ctx->desc_done, stream));
    return ICEDEM_OK;
}

}  // extern "C"

// Kernel timing for icedem_frames_profiled: every launch gets a start/stop
// event pair recorded by its own dispatch (hipExtLaunchKernelGGL), so a pair
// spans exactly that kernel's execution.  Category: 0 unused (a separate
// near-test kernel, no longer launched), 1 k_force (with sub-step cell
// rebuilds), 2 k_post, 3 k_cull.
constexpr int NPROF = 4;
struct LaunchTimer {
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pairs;
    int slot(int cat, hipEvent_t& a, hipEvent_t& b) {
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return -1;
        pairs.push_back({cat, {a, b}});
        return 0;
    }
    ~LaunchTimer() {
        for (auto& p : pairs) { (void)hipEventDestroy(p.second.first); (void)hipEventDestroy(p.second.second); }
    }
};

template <typename F, typename... Args>
static int launch(LaunchTimer* lt, int cat, F kernel, dim3 grid, dim3 block, hipStream_t stream, Args... args) {
    if (!lt) {
        hipLaunchKernelGGL(kernel, grid, block, 0, stream, args...);
        return ICEDEM_OK;
    }
    hipEvent_t a = nullptr, b = nullptr;
    if (lt->slot(cat, a, b)) return ICEDEM_E_HIP;
    hipExtLaunchKernelGGL(kernel, grid, block, 0, stream, a, b, 0, args...);
    return ICEDEM_OK;
}

static int launch_frames(IceDemCtx* ctx, int K, hipStream_t stream, LaunchTimer* lt = nullptr) {
    const DevView& H = ctx->hv;
    const DevView* V = ctx->dv;
    const CallDesc* D = ctx->dd;
    const int nblk = H.nblk;
    const int fblk = force_chunks(H.n, H.lanes);
    const int nextra = TAIL_GROUPS;
    const int steps = H.p.n_ice_steps;
    // post grid: element/cell/floe blocks, then bond blocks
    const int nbe = (std::max(H.n, std::max(std::max(H.G, H.G2), H.nf)) + BLOCK - 1) / BLOCK;
    const int nbpost = nbe + (H.nb + BLOCK - 1) / BLOCK;
    const int gcl = (std::max(H.G, H.G2) + BLOCK - 1) / BLOCK;
    // the force-independent part of each frame's ship step is formed by one
    // extra workgroup of the kernel before the frame's force pass (not with
    // the ice brake, whose setpoint depends on the frame's ice force)
    const int sb = (H.p.ship_mode == 1 || H.p.ice_brake_threshold <= 0.0f) ? 1 : 0;
    // k_post's near workgroups (a multiple of 8): one wave per candidate,
    // at least one candidate per 64 elements' worth of waves
    const int nnw = std::max(8, std::min(64, ((nblk + 7) / 8) * 8));
    int rc = launch(lt, 3, k_cull, dim3(nblk + sb), dim3(BLOCK), stream, V, D, sb);
    for (int k = 0; k < K && rc == ICEDEM_OK; ++k) {
        const int last = (k == K - 1) ? 1 : 0;
        const int set = k & 1;
        for (int s = 0; s < steps; ++s) {
            if (s > 0) {
                // rebuild the cell buckets for the sub-step's positions
                const bool odd = (s % 2 == 1);
                const float* sp = odd ? (const float*)H.nxt_pos : (const float*)nullptr;
                const float* sv = odd ? (const float*)H.nxt_vel : (const float*)nullptr;
                rc |= launch(lt, 1, k_clear_cells, dim3(gcl), dim3(BLOCK), stream, V, set);
                rc |= launch(lt, 1, k_cellkey, dim3(nblk), dim3(BLOCK), stream, V, D, sp, sv, odd ? 0 : 1, set);
            }
            const int extra = (s == 0) ? nextra : 0;
            rc |= launch(lt, 1, H.lanes == 8 ? k_force<8> : k_force<16>, dim3(fblk + extra), dim3(BLOCK), stream,
                         V, D, k, s, (s % 2 == 0) ? 1 : 0, 0, set, last, extra,
                         (const float*)nullptr, (float*)nullptr, (float*)nullptr, sb);
        }
        const int psb = last ? 0 : sb;
        const int pnw = last ? 0 : nnw;
        rc |= launch(lt, 2, k_post, dim3(pnw + nbpost + psb), dim3(BLOCK), stream, V, D, k,
                     (steps % 2 == 1) ? 1 : 0, last, set, last ? 0 : 1, nbe, psb, pnw);
    }
    if (rc) { set_err(ctx, "event creation failed"); return ICEDEM_E_HIP; }
    HIP_TRY(hipGetLastError());
    return ICEDEM_OK;
}

extern "C" {

int icedem_frames(IceDemCtx* ctx, const IceDemState* state, const IceDemInputs* in, int32_t K,
                  const IceDemHistory* hist, const IceDemDiag* diag, int32_t use_graph, void* stream_) {
    if (!ctx || !in || K < 0) { set_err(ctx, "null argument"); return ICEDEM_E_INVALID; }
    if (K == 0) return ICEDEM_OK;
    int rc = validate_state(ctx, state);
    if (rc) return rc;
    if (!in->d_setpoints) { set_err(ctx, "setpoints required"); return ICEDEM_E_INVALID; }
    hipStream_t stream = (hipStream_t)stream_;
    HIP_TRY(hipSetDevice(ctx->device));
    CallDesc cd;
    memset(&cd, 0, sizeof(cd));
    cd.s = *state;
    cd.in = *in;
    if (hist) cd.h = *hist;
    if (diag) cd.dg = *diag;
    rc = put_desc(ctx, cd, stream);
    if (rc) return rc;
    if (!use_graph) return launch_frames(ctx, K, stream);
    auto it = ctx->graphs.find(K);
    if (it == ctx->graphs.end()) {
        hipStream_t cap;
        HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        HIP_TRY(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        rc = launch_frames(ctx, K, cap);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(cap, &g);
        (void)hipStreamDestroy(cap);
        if (rc) return rc;
        if (e2 != hipSuccess) { set_err(ctx, std::string("capture: ") + hipGetErrorString(e2)); return ICEDEM_E_HIP; }
        hipGraphExec_t ge = nullptr;
        HIP_TRY(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        ctx->graph_defs[K] = g;
        ctx->graphs[K] = ge;
        it = ctx->graphs.find(K);
    }
    HIP_TRY(hipGraphLaunch(it->second, stream));
    return ICEDEM_OK;
}

int icedem_list_stats(IceDemCtx* ctx, const uint8_t* d_active, int32_t* h_out) {
    if (!ctx || !h_out) { set_err(ctx, "null argument"); return ICEDEM_E_INVALID; }
    h_out[0] = 0; h_out[1] = 0; h_out[2] = ctx->hv.nl_on; h_out[3] = 0; h_out[4] = 0; h_out[5] = 0;
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<uint8_t> act;
    if (d_active) {
        act.resize(ctx->hv.n);
        HIP_TRY(hipMemcpy(act.data(), d_active, act.size(), hipMemcpyDeviceToHost));
        int64_t na = 0;
        for (uint8_t a : act) na += a != 0;
        h_out[5] = (int32_t)na;
    }
    if (!ctx->hv.nl_on) return ICEDEM_OK;
    uint32_t w[2];
    HIP_TRY(hipMemcpy(w, ctx->hv.nlctl + 3 * NLSUB, sizeof(w), hipMemcpyDeviceToHost));
    h_out[0] = (int32_t)w[1];
    h_out[1] = (int32_t)w[0];
    std::vector<int32_t> cnt(ctx->hv.n);
    HIP_TRY(hipMemcpy(cnt.data(), ctx->hv.nl_cnt, sizeof(int32_t) * cnt.size(), hipMemcpyDeviceToHost));
    int64_t tot = 0, tot_act = 0;
    for (size_t e = 0; e < cnt.size(); ++e) {
        const int64_t c = (cnt[e] >= 0 && cnt[e] <= NLW) ? cnt[e] : 0;
        tot += c;
        if (d_active && act[e]) tot_act += c;
    }
    h_out[3] = (int32_t)std::min<int64_t>(tot, INT32_MAX);
    h_out[4] = (int32_t)std::min<int64_t>(d_active ? tot_act : tot, INT32_MAX);
    return ICEDEM_OK;
}

int icedem_frames_profiled(IceDemCtx* ctx, const IceDemState* state, const IceDemInputs* in, int32_t K,
                           const IceDemHistory* hist, float* h_kernel_ms, void* stream_) {
    if (!ctx || !in || !h_kernel_ms || K < 0) { set_err(ctx, "null argument"); return ICEDEM_E_INVALID; }
    for (int i = 0; i < NPROF; ++i) h_kernel_ms[i] = 0.0f;
    if (K == 0) return ICEDEM_OK;
    int rc = validate_state(ctx, state);
    if (rc) return rc;
    if (!in->d_setpoints) { set_err(ctx, "setpoints required"); return ICEDEM_E_INVALID; }
    hipStream_t stream = (hipStream_t)stream_;
    HIP_TRY(hipSetDevice(ctx->device));
    CallDesc cd;
    memset(&cd, 0, sizeof(cd));
    cd.s = *state;
    cd.in = *in;
    if (hist) cd.h = *hist;
    rc = put_desc(ctx, cd, stream);
    if (rc) return rc;
    LaunchTimer lt;
    rc = launch_frames(ctx, K, stream, &lt);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(stream));
    float acc[NPROF] = {0, 0, 0, 0};
    for (auto& p : lt.pairs) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, p.second.first, p.second.second));
        acc[p.first] += ms;
    }
    for (int i = 0; i < NPROF; ++i) h_kernel_ms[i] = acc[i];
    return ICEDEM_OK;
}

struct IceDemBatch {
    int device = 0;
    int n = 0;
    int mode = ICEDEM_BATCH_FRAMES;
    DevView* dviews = nullptr;   // (n) copies of the contexts' device views
    CallDesc* ddesc = nullptr;   // (n)
    CallDesc* hdesc = nullptr;   // pinned staging
    hipEvent_t desc_done = nullptr;
    // workgroup maps of the batched frame kernels: (instance, workgroup of
    // the instance, the instance's workgroup count, the instance's nextra)
    int4* map_cull = nullptr; int4* map_force = nullptr; int4* map_post = nullptr;
    int n_cull = 0, n_force = 0, n_post = 0;
    int lanes = GRP;             // lanes per element of kb_force
    std::map<std::pair<int, int>, hipGraphExec_t> graphs;   // (frames, mode)
    std::map<std::pair<int, int>, hipGraph_t> graph_defs;
    std::string err;
};

// Hardware workgroup w of a launch runs on XCD w % 8.  Re-deal a batch map
// (each instance's entries contiguous) so that every instance's workgroups
// run on one XCD, instances spread over the XCDs by size: an instance's
// bucket records and partner state then stay in one L2.  With interleave,
// an XCD takes its instances' first entries, then their second ones, and so
// on (no instance's workgroups all queue behind the others': a patch's
// first chunk, where the hull enters, starts with the kernel).  Entries are
// self-contained (instance, workgroup, count, extra), so the order changes
// no result.
static std::vector<int4> xcd_deal(const std::vector<int4>& m, bool interleave = false) {
    constexpr int NX = 8;
    std::vector<std::pair<size_t, size_t>> rng;   // (first entry, entries) per instance
    for (size_t a = 0; a < m.size();) {
        size_t z = a;
        while (z < m.size() && m[z].x == m[a].x) ++z;
        rng.push_back({a, z - a});
        a = z;
    }
    std::vector<size_t> ord(rng.size());
    for (size_t i = 0; i < ord.size(); ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](size_t p, size_t q) { return rng[p].second > rng[q].second; });
    std::vector<std::vector<int4>> L(NX);
    std::vector<std::vector<size_t>> own(NX);   // instances (ranges) of each XCD
    for (size_t r : ord) {
        int x = 0;
        for (int y = 1; y < NX; ++y) if (L[y].size() < L[x].size()) x = y;
        L[x].insert(L[x].end(), m.begin() + rng[r].first, m.begin() + rng[r].first + rng[r].second);
        own[x].push_back(r);
    }
    if (interleave) {
        for (int x = 0; x < NX; ++x) {
            L[x].clear();
            for (size_t j = 0;; ++j) {
                bool any = false;
                for (size_t r : own[x])
                    if (j < rng[r].second) { L[x].push_back(m[rng[r].first + j]); any = true; }
                if (!any) break;
            }
        }
    }
    std::vector<int4> out;
    out.reserve(m.size());
    std::vector<size_t> pos(NX, 0);
    for (size_t w = 0; w < m.size(); ++w) {
        int x = (int)(w % NX);
        if (pos[x] >= L[x].size()) {   // this XCD's list is spent: the longest remaining one
            for (int y = 0; y < NX; ++y)
                if (L[y].size() - pos[y] > L[x].size() - pos[x]) x = y;
        }
        out.push_back(L[x][pos[x]++]);
    }
    return out;
}

// Tail entries t placed before the dealt element entries m: position p of
// the result runs on XCD p % 8, and a tail goes where its instance's element
// workgroups run while that XCD has tails left (else on the XCD with the most
// tails left).
static std::vector<int4> tails_first(const std::vector<int4>& t, const std::vector<int4>& m, int n_inst) {
    constexpr int NX = 8;
    const size_t T = t.size();
    std::vector<int> xcd(n_inst, -1);
    for (size_t w = 0; w < m.size(); ++w)
        if (xcd[m[w].x] < 0) xcd[m[w].x] = (int)((T + w) % NX);
    std::vector<std::vector<int4>> q(NX);
    for (const int4& e : t) q[xcd[e.x] >= 0 ? xcd[e.x] : 0].push_back(e);
    std::vector<size_t> pos(NX, 0);
    std::vector<int4> out;
    out.reserve(T + m.size());
    for (size_t p = 0; p < T; ++p) {
        int x = (int)(p % NX);
        if (pos[x] >= q[x].size()) {
            for (int y = 0; y < NX; ++y)
                if (q[y].size() - pos[y] > q[x].size() - pos[x]) x = y;
        }
        out.push_back(q[x][pos[x]++]);
    }
    out.insert(out.end(), m.begin(), m.end());
    return out;
}

static void batch_free(IceDemBatch* b) {
    for (auto& g : b->graphs) (void)hipGraphExecDestroy(g.second);
    for (auto& g : b->graph_defs) (void)hipGraphDestroy(g.second);
    if (b->dviews) (void)hipFree(b->dviews);
    if (b->ddesc) (void)hipFree(b->ddesc);
    if (b->hdesc) (void)hipHostFree(b->hdesc);
    if (b->desc_done) (void)hipEventDestroy(b->desc_done);
    for (int4* m : {b->map_cull, b->map_force, b->map_post})
        if (m) (void)hipFree(m);
    delete b;
}

int icedem_batch_create(IceDemCtx* const* ctxs, int32_t n_inst, IceDemBatch** out) {
    if (!ctxs || n_inst < 1 || !out) { set_err(nullptr, "bad argument"); return ICEDEM_E_INVALID; }
    *out = nullptr;
    for (int i = 0; i < n_inst; ++i) {
        if (!ctxs[i]) { set_err(nullptr, "null context"); return ICEDEM_E_INVALID; }
        const DevView& H = ctxs[i]->hv;
        if (H.nbpad > 256 || H.p.n_ice_steps != 1 || ctxs[i]->device != ctxs[0]->device) {
            set_err(nullptr, "batched scenes need n <= 65536, one ice sub-step and one device");
            return ICEDEM_E_LIMIT;
        }
    }
    {   // each instance works in its own context's scratch
        std::vector<const IceDemCtx*> v(ctxs, ctxs + n_inst);
        std::sort(v.begin(), v.end());
        if (std::adjacent_find(v.begin(), v.end()) != v.end()) {
            set_err(nullptr, "a context appears twice in the batch"); return ICEDEM_E_INVALID;
        }
    }
    IceDemBatch* b = new IceDemBatch();
    b->device = ctxs[0]->device;
    b->n = n_inst;
    {
        int64_t total = 0;
        for (int i = 0; i < n_inst; ++i) total += ctxs[i]->hv.n;
        b->lanes = resolve_lanes(ctxs[0]->hv.p.force_lanes, total);
    }
    std::vector<int4> mc, mn, mt, mf, mp;
    for (int i = 0; i < n_inst; ++i) {
        const DevView& H = ctxs[i]->hv;
        const int fblk = force_chunks(H.n, b->lanes);
        const int nextra = TAIL_GROUPS;
        // kb_post's near workgroups of the instance: one wavefront per 64
        // elements and at least 8 (each wave loops over the candidates);
        // on the 256-patch batch 29.6 us per frame against 30.2 with one
        // wavefront per 16 elements, 31.5 with one per 8, 30.5 with a floor
        // of 4 waves
        const int nnear = std::max(2, std::min(64, (H.n + 255) / 256));
        const int nbe = (std::max(H.n, std::max(std::max(H.G, H.G2), H.nf)) + BLOCK - 1) / BLOCK;
        const int npost = nbe + (H.nb + BLOCK - 1) / BLOCK;
        for (int l = 0; l < H.nblk; ++l) mc.push_back(make_int4(i, l, H.nblk, 0));
        for (int l = 0; l < nnear; ++l) mn.push_back(make_int4(i, l, nnear, -1));
        for (int l = 0; l < nextra; ++l) mt.push_back(make_int4(i, l, nextra + fblk, nextra));
        for (int l = nextra; l < nextra + fblk; ++l) mf.push_back(make_int4(i, l, nextra + fblk, nextra));
        for (int l = 0; l < npost; ++l) mp.push_back(make_int4(i, l, npost, nbe));
    }
    // kb_post: every instance's near workgroups first, then the post blocks
    mp = xcd_deal(mp, true);
    mp.insert(mp.begin(), mn.begin(), mn.end());
    // kb_force only: 256 patches 64.3 -> 62.2 us per frame; dealing kb_post
    // the same way gained less (62.4)
    mf = xcd_deal(mf, true);
    // the instances' tail workgroups (ship forces, ship step, eligibility: a
    // chain of dependent round trips) come first, each on its instance's XCD
    // where the counts allow, so every chain starts with the kernel instead
    // of behind the element blocks dealt before it
    mf = tails_first(mt, mf, n_inst);
    b->n_cull = (int)mc.size(); b->n_force = (int)mf.size(); b->n_post = (int)mp.size();
    hipError_t e = hipSetDevice(b->device);
    std::vector<DevView> hv(n_inst);
    for (int i = 0; i < n_inst; ++i) hv[i] = ctxs[i]->hv;
    auto upmap = [&](int4** d, const std::vector<int4>& h) {
        if (e == hipSuccess) e = hipMalloc((void**)d, sizeof(int4) * h.size());
        if (e == hipSuccess) e = hipMemcpy(*d, h.data(), sizeof(int4) * h.size(), hipMemcpyHostToDevice);
    };
    if (e == hipSuccess) e = hipMalloc((void**)&b->dviews, sizeof(DevView) * n_inst);
    if (e == hipSuccess) e = hipMalloc((void**)&b->ddesc, sizeof(CallDesc) * n_inst);
    if (e == hipSuccess) e = hipMemcpy(b->dviews, hv.data(), sizeof(DevView) * n_inst, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipHostMalloc((void**)&b->hdesc, sizeof(CallDesc) * n_inst, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->desc_done, hipEventDisableTiming);
    upmap(&b->map_cull, mc); upmap(&b->map_force, mf); upmap(&b->map_post, mp);
    if (e != hipSuccess) {
        set_err(nullptr, std::string("icedem_batch_create: ") + hipGetErrorString(e));
        batch_free(b);
        return ICEDEM_E_HIP;
    }
    memset(b->hdesc, 0, sizeof(CallDesc) * n_inst);
    *out = b;
    return ICEDEM_OK;
}

int icedem_batch_destroy(IceDemBatch* b) {
    if (!b) return ICEDEM_OK;
    (void)hipSetDevice(b->device);
    batch_free(b);
    return ICEDEM_OK;
}

int icedem_batch_mode(IceDemBatch* b, int32_t mode) {
    if (!b || (mode != ICEDEM_BATCH_FRAMES && mode != ICEDEM_BATCH_SCENES)) {
        set_err(nullptr, "bad batch mode"); return ICEDEM_E_INVALID;
    }
    b->mode = mode;
    return ICEDEM_OK;
}

// the launch sequence of `frames` frames over every instance
static void batch_launch(IceDemBatch* b, int frames, hipStream_t stream) {
    if (b->mode == ICEDEM_BATCH_SCENES) {
        hipLaunchKernelGGL(k_scene_rollouts, dim3(b->n), dim3(BLOCK), 0, stream, b->dviews, b->ddesc, frames);
        return;
    }
    hipLaunchKernelGGL(kb_cull, dim3(b->n_cull), dim3(BLOCK), 0, stream, b->dviews, b->ddesc,
                       (const int4*)b->map_cull);
    for (int k = 0; k < frames; ++k) {
        const int last = (k == frames - 1) ? 1 : 0;
        const int set = k & 1;
        hipLaunchKernelGGL(b->lanes == 8 ? kb_force<8> : kb_force<16>, dim3(b->n_force), dim3(BLOCK), 0, stream, b->dviews, b->ddesc,
                           (const int4*)b->map_force, k, last, set);
        hipLaunchKernelGGL(kb_post, dim3(b->n_post), dim3(BLOCK), 0, stream, b->dviews, b->ddesc,
                           (const int4*)b->map_post, k, last, set);
    }
}

int icedem_batch_rollouts(IceDemBatch* b, const IceDemState* states, const IceDemInputs* inputs, int32_t frames,
                          const IceDemDiag* diags, int32_t use_graph, void* stream_) {
    IceDemCtx* ctx = nullptr;
    if (!b || !states || !inputs || frames < 0) { set_err(nullptr, "bad argument"); return ICEDEM_E_INVALID; }
    if (frames == 0) return ICEDEM_OK;
    hipStream_t stream = (hipStream_t)stream_;
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipEventSynchronize(b->desc_done));
    for (int i = 0; i < b->n; ++i) {
        const IceDemState& s = states[i];
        if (!s.d_pos || !s.d_vel || !s.d_rot || !s.d_omega || !s.d_sub || !s.d_active || !s.d_ship) {
            set_err(nullptr, "batch state has a null pointer"); return ICEDEM_E_INVALID;
        }
        CallDesc& cd = b->hdesc[i];
        memset(&cd, 0, sizeof(cd));
        cd.s = s;
        cd.in = inputs[i];
        if (diags) cd.dg = diags[i];
    }
    HIP_TRY(hipMemcpyAsync(b->ddesc, b->hdesc, sizeof(CallDesc) * b->n, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipEventRecord(b->desc_done, stream));
    if (!use_graph) {
        batch_launch(b, frames, stream);
        HIP_TRY(hipGetLastError());
        return ICEDEM_OK;
    }
    const std::pair<int, int> key(frames, b->mode);
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
        hipStream_t cap;
        HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        HIP_TRY(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        batch_launch(b, frames, cap);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(cap, &g);
        (void)hipStreamDestroy(cap);
        if (e2 != hipSuccess) { set_err(nullptr, std::string("capture: ") + hipGetErrorString(e2)); return ICEDEM_E_HIP; }
        hipGraphExec_t ge = nullptr;
        HIP_TRY(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        b->graph_defs[key] = g;
        b->graphs[key] = ge;
        it = b->graphs.find(key);
    }
    HIP_TRY(hipGraphLaunch(it->second, stream));
    return ICEDEM_OK;
}

int icedem_sync(IceDemCtx* ctx, void* stream) {
    if (!ctx) return ICEDEM_E_INVALID;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    int32_t e[4];
    HIP_TRY(hipMemcpy(e, ctx->hv.err, sizeof(e), hipMemcpyDeviceToHost));
    if (e[0] != 0) { set_err(ctx, "kernel flagged error " + std::to_string(e[0])); return ICEDEM_E_DEVICE; }
    return ICEDEM_OK;
}

int icedem_substep_forces(IceDemCtx* ctx, const IceDemState* state, const float* d_ship_forces,
                          const float* d_v_water, int64_t vw_stride_c, int64_t vw_stride_e,
                          float* d_forces, float* d_moments, void* stream_) {
    if (!ctx || !d_ship_forces || !d_forces || !d_moments) { set_err(ctx, "null argument"); return ICEDEM_E_INVALID; }
    int rc = validate_state(ctx, state);
    if (rc) return rc;
    hipStream_t stream = (hipStream_t)stream_;
    HIP_TRY(hipSetDevice(ctx->device));
    CallDesc cd;
    memset(&cd, 0, sizeof(cd));
    cd.s = *state;
    cd.in.d_v_water = d_v_water;
    cd.in.vw_stride_c = vw_stride_c;
    cd.in.vw_stride_e = vw_stride_e;
    rc = put_desc(ctx, cd, stream);
    if (rc) return rc;
    const DevView& H = ctx->hv;
    const int fblk = force_chunks(H.n, H.lanes);
    hipLaunchKernelGGL(k_cellkey, dim3(H.nblk), dim3(BLOCK), 0, stream, ctx->dv, ctx->dd, (const float*)nullptr,
                       (const float*)nullptr, 1, 0);
    hipLaunchKernelGGL(H.lanes == 8 ? k_force<8> : k_force<16>, dim3(fblk), dim3(BLOCK), 0, stream, ctx->dv, ctx->dd, 0, 0, 1, 1, 0, 0, 0,
                       d_ship_forces, d_forces, d_moments, 0);
    hipLaunchKernelGGL(k_clear_cells, dim3((std::max(H.G, H.G2) + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, stream, ctx->dv, 0);
    HIP_TRY(hipGetLastError());
    return ICEDEM_OK;
}

int icedem_contact_pairs(IceDemCtx* ctx, const float* d_pos, int32_t* d_pair_i, int32_t* d_pair_j,
                         int64_t max_pairs, int64_t* h_n_pairs, void* stream_) {
    if (!ctx || !d_pos || !h_n_pairs || max_pairs < 0 || (max_pairs > 0 && (!d_pair_i || !d_pair_j))) {
        set_err(ctx, "bad argument"); return ICEDEM_E_INVALID;
    }
    hipStream_t stream = (hipStream_t)stream_;
    HIP_TRY(hipSetDevice(ctx->device));
    const DevView& H = ctx->hv;
    int32_t* cnt = nullptr;
    int32_t* off = nullptr;
    HIP_TRY(hipMallocAsync((void**)&cnt, sizeof(int32_t) * (H.n + 1), stream));
    HIP_TRY(hipMallocAsync((void**)&off, sizeof(int32_t) * (H.n + 1), stream));
    hipLaunchKernelGGL(k_cellkey, dim3(H.nblk), dim3(BLOCK), 0, stream, ctx->dv, ctx->dd, d_pos, d_pos, 0, 0);
    hipLaunchKernelGGL(k_pair_count, dim3(H.nblk), dim3(BLOCK), 0, stream, ctx->dv, d_pos, cnt);
    hipLaunchKernelGGL(k_excl_scan, dim3(1), dim3(1024), 0, stream, H.n, cnt, off);
    hipLaunchKernelGGL(k_pair_emit, dim3(H.nblk), dim3(BLOCK), 0, stream, ctx->dv, d_pos, off, max_pairs,
                       d_pair_i, d_pair_j);
    hipLaunchKernelGGL(k_clear_cells, dim3((std::max(H.G, H.G2) + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, stream, ctx->dv, 0);
    HIP_TRY(hipGetLastError());
    int32_t total = 0;
    HIP_TRY(hipMemcpyAsync(&total, off + H.n, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipFreeAsync(cnt, stream));
    HIP_TRY(hipFreeAsync(off, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    *h_n_pairs = total;
    return ICEDEM_OK;
}

int icedem_collisions_combined(const float* d_local_polys, const uint8_t* d_poly_mask, const float* d_pos,
                               const float* d_rot, int32_t n, int32_t max_v, float ship_x, float ship_y,
                               float ship_heading, float half_len, float half_beam, const float* d_hull_local,
                               int32_t n_hull, int32_t n_edge_samples, const float* h_edge_t,
                               const float* h_edge_omt, float* d_pen, float* d_nx, float* d_ny, float* d_cx,
                               float* d_cy, void* stream_) {
    IceDemCtx* ctx = nullptr;
    if (n < 0 || max_v < 1 || max_v > ICEDEM_MAX_VERTS || n_hull < 1 || n_edge_samples < 0 ||
        n_edge_samples > ICEDEM_MAX_EDGE_SAMPLES || (n_edge_samples > 0 && (!h_edge_t || !h_edge_omt))) {
        set_err(nullptr, "bad argument"); return ICEDEM_E_INVALID;
    }
    if (n == 0) return ICEDEM_OK;
    hipStream_t stream = (hipStream_t)stream_;
    IceDemParams prm;
    memset(&prm, 0, sizeof(prm));
    prm.half_len = half_len; prm.half_beam = half_beam; prm.n_edge_samples = n_edge_samples;
    for (int i = 0; i < n_edge_samples; ++i) { prm.edge_t[i] = h_edge_t[i]; prm.edge_omt[i] = h_edge_omt[i]; }
    // hull cull radius from the hull points and the analytic box
    std::vector<float> hl(2 * (size_t)n_hull);
    HIP_TRY(hipMemcpy(hl.data(), d_hull_local, sizeof(float) * hl.size(), hipMemcpyDeviceToHost));
    double R = std::sqrt((double)half_len * half_len + (double)half_beam * half_beam);
    for (int s = 0; s < n_hull; ++s) R = std::max(R, std::hypot((double)hl[2 * s], (double)hl[2 * s + 1]));
    prm.ship_bound_radius = (float)(R * 1.0001 + 1e-3);
    float* bound = nullptr;
    HIP_TRY(hipMallocAsync((void**)&bound, sizeof(float) * n, stream));
    hipLaunchKernelGGL(k_bound_radius, dim3((n + 255) / 256), dim3(256), 0, stream,
                       (const float2*)d_local_polys, n, max_v, bound);
    hipLaunchKernelGGL(k_collisions_leaf, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, stream,
                       (const float2*)d_local_polys, d_poly_mask, d_pos, d_rot, (const float*)bound, n, max_v,
                       ship_x, ship_y, ship_heading, (const float2*)d_hull_local, n_hull, prm,
                       d_pen, d_nx, d_ny, d_cx, d_cy);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipFreeAsync(bound, stream));
    return ICEDEM_OK;
}

int icedem_bond_damage(int32_t n_bonds, const float* d_delta_n, const float* d_delta_s, const float* d_sigma_n,
                       const float* d_sigma_t, const float* d_sigma_bar, float C, float mu_b, float G_c_I,
                       float G_c_II, float eta_bk, const float* d_damage, const float* d_dm_max, const float* d_dm_0,
                       const uint8_t* d_intact, float* d_new_damage, float* d_new_dm_max, float* d_new_dm_0,
                       uint8_t* d_new_intact, void* stream_) {
    IceDemCtx* ctx = nullptr;
    if (n_bonds < 0) { set_err(nullptr, "bad argument"); return ICEDEM_E_INVALID; }
    if (n_bonds == 0) return ICEDEM_OK;
    hipLaunchKernelGGL(k_bond_damage_leaf, dim3((n_bonds + 255) / 256), dim3(256), 0, (hipStream_t)stream_,
                       n_bonds, d_delta_n, d_delta_s, d_sigma_n, d_sigma_t, d_sigma_bar, C, mu_b, G_c_I, G_c_II,
                       eta_bk, d_damage, d_dm_max, d_dm_0, d_intact, d_new_damage, d_new_dm_max, d_new_dm_0,
                       d_new_intact);
    HIP_TRY(hipGetLastError());
    return ICEDEM_OK;
}

int icedem_connected_components(int32_t n, int32_t n_bonds, const int32_t* d_bond_i, const int32_t* d_bond_j,
                                const uint8_t* d_intact, int32_t n_iter, int32_t* d_labels, void* stream_) {
    IceDemCtx* ctx = nullptr;
    if (n < 0 || n_bonds < 0 || n_iter < 0) { set_err(nullptr, "bad argument"); return ICEDEM_E_INVALID; }
    if (n == 0) return ICEDEM_OK;
    hipStream_t stream = (hipStream_t)stream_;
    int32_t* tmp = nullptr;
    HIP_TRY(hipMallocAsync((void**)&tmp, sizeof(int32_t) * n, stream));
    std::vector<int32_t> iota(n);
    for (int i = 0; i < n; ++i) iota[i] = i;
    HIP_TRY(hipMemcpyAsync(d_labels, iota.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, stream));
    const int gb = (n + 255) / 256, gbb = (std::max(n_bonds, 1) + 255) / 256;
    for (int it = 0; it < n_iter; ++it) {
        hipLaunchKernelGGL(k_cc_copy, dim3(gb), dim3(256), 0, stream, n, (const int32_t*)d_labels, tmp);
        if (n_bonds > 0)
            hipLaunchKernelGGL(k_cc_sweep, dim3(gbb), dim3(256), 0, stream, n, n_bonds, d_bond_i, d_bond_j,
                               d_intact, (const int32_t*)d_labels, tmp);
        hipLaunchKernelGGL(k_cc_copy, dim3(gb), dim3(256), 0, stream, n, (const int32_t*)tmp, d_labels);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(stream));
    HIP_TRY(hipFreeAsync(tmp, stream));
    return ICEDEM_OK;
}

}  // extern "C"

// ── connected components without allocation ──────────────────────────────────
// The label scratch of icedem_connected_components_noalloc: ICEDEM_CC_NOALLOC_MAX
// int32 (4 MiB), one copy on every GPU the library runs on, held until the
// process ends.  Host code never passes `scratch` itself to a kernel (in host
// code the name is a stand-in, not a device address): each device's address
// comes from hipGetSymbolAddress, once per device, with that device current.
static __device__ __align__(256) int32_t scratch[ICEDEM_CC_NOALLOC_MAX];

constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static int32_t* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// this device's address of `scratch` (looked up on the device's first call)
static int cc_scratch(int32_t** out) {
    IceDemCtx* ctx = nullptr;
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) { set_err(nullptr, "device index out of range"); return ICEDEM_E_LIMIT; }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_err(nullptr, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return ICEDEM_E_HIP;
        }
        g_scratch_ptr[dev] = (int32_t*)p;
    }
    *out = g_scratch_ptr[dev];
    return ICEDEM_OK;
}

extern "C" {

int icedem_connected_components_noalloc(int32_t n, int32_t n_bonds, const int32_t* d_bond_i,
                                        const int32_t* d_bond_j, const uint8_t* d_intact, int32_t n_iter,
                                        int32_t* d_labels, void* stream_) {
    IceDemCtx* ctx = nullptr;
    if (n < 0 || n_bonds < 0 || n_iter < 0) { set_err(nullptr, "bad argument"); return ICEDEM_E_INVALID; }
    if (n > ICEDEM_CC_NOALLOC_MAX) {
        set_err(nullptr, "icedem_connected_components_noalloc: n = " + std::to_string(n) + " exceeds the scratch "
                         "limit of " + std::to_string(ICEDEM_CC_NOALLOC_MAX) + " elements");
        return ICEDEM_E_LIMIT;
    }
    if (n == 0) return ICEDEM_OK;
    if (!d_labels || (n_bonds > 0 && (!d_bond_i || !d_bond_j || !d_intact))) {
        set_err(nullptr, "null argument"); return ICEDEM_E_INVALID;
    }
    int32_t* tmp = nullptr;
    int rc = cc_scratch(&tmp);
    if (rc) return rc;
    hipStream_t stream = (hipStream_t)stream_;
    // every sweep overwrites the scratch's first n entries before reading
    // them, so nothing is cleared; what a previous call left there is unused
    const int gb = (n + 255) / 256, gbb = (std::max(n_bonds, 1) + 255) / 256;
    hipLaunchKernelGGL(k_cc_init, dim3(gb), dim3(256), 0, stream, n, d_labels);
    for (int it = 0; it < n_iter; ++it) {
        hipLaunchKernelGGL(k_cc_copy, dim3(gb), dim3(256), 0, stream, n, (const int32_t*)d_labels, tmp);
        if (n_bonds > 0)
            hipLaunchKernelGGL(k_cc_sweep, dim3(gbb), dim3(256), 0, stream, n, n_bonds, d_bond_i, d_bond_j,
                               d_intact, (const int32_t*)d_labels, tmp);
        hipLaunchKernelGGL(k_cc_copy, dim3(gb), dim3(256), 0, stream, n, (const int32_t*)tmp, d_labels);
    }
    HIP_TRY(hipGetLastError());
    return ICEDEM_OK;
}

int icedem_component_com_velocities(int32_t n, const int32_t* d_labels, const float* d_masses, const float* d_vel,
                                    float* d_out, void* stream_) {
    IceDemCtx* ctx = nullptr;
    if (n < 0) { set_err(nullptr, "bad argument"); return ICEDEM_E_INVALID; }
    if (n == 0) return ICEDEM_OK;
    hipStream_t stream = (hipStream_t)stream_;
    int32_t *count, *start, *fill, *order;
    float *tm, *tx, *ty;
    HIP_TRY(hipMallocAsync((void**)&count, sizeof(int32_t) * (n + 1), stream));
    HIP_TRY(hipMallocAsync((void**)&start, sizeof(int32_t) * (n + 1), stream));
    HIP_TRY(hipMallocAsync((void**)&fill, sizeof(int32_t) * (n + 1), stream));
    HIP_TRY(hipMallocAsync((void**)&order, sizeof(int32_t) * n, stream));
    HIP_TRY(hipMallocAsync((void**)&tm, sizeof(float) * n, stream));
    HIP_TRY(hipMallocAsync((void**)&tx, sizeof(float) * n, stream));
    HIP_TRY(hipMallocAsync((void**)&ty, sizeof(float) * n, stream));
    HIP_TRY(hipMemsetAsync(count, 0, sizeof(int32_t) * (n + 1), stream));
    HIP_TRY(hipMemsetAsync(fill, 0, sizeof(int32_t) * (n + 1), stream));
    const int gb = (n + 255) / 256;
    hipLaunchKernelGGL(k_iota_count, dim3(gb), dim3(256), 0, stream, n, d_labels, count);
    hipLaunchKernelGGL(k_excl_scan, dim3(1), dim3(1024), 0, stream, n, (const int32_t*)count, start);
    hipLaunchKernelGGL(k_stable_place, dim3(1), dim3(1024), 0, stream, n, d_labels, (const int32_t*)start, fill, order);
    hipLaunchKernelGGL(k_com_sum, dim3(gb), dim3(256), 0, stream, n, (const int32_t*)start, (const int32_t*)order,
                       d_masses, d_vel, tm, tx, ty);
    hipLaunchKernelGGL(k_com_out, dim3(gb), dim3(256), 0, stream, n, d_labels, (const float*)tm, (const float*)tx,
                       (const float*)ty, d_out);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipFreeAsync(count, stream)); HIP_TRY(hipFreeAsync(start, stream)); HIP_TRY(hipFreeAsync(fill, stream));
    HIP_TRY(hipFreeAsync(order, stream)); HIP_TRY(hipFreeAsync(tm, stream)); HIP_TRY(hipFreeAsync(tx, stream));
    HIP_TRY(hipFreeAsync(ty, stream));
    return ICEDEM_OK;
}

int icedem_ship_rollouts(IceDemCtx* ctx, const float* d_ship0, int32_t ship0_stride, const float* d_controls, int32_t S,
                         int32_t n_knots, int32_t hold, int32_t frames, const float* d_costmap, int32_t rows,
                         int32_t cols, float x0, float y0, float res, float oob_penalty, float* d_out,
                         int32_t snap_every, int32_t n_snap, float* d_snap, void* stream_) {
    if (!ctx || !d_ship0 || !d_controls || !d_out) { set_err(ctx, "null argument"); return ICEDEM_E_INVALID; }
    if (S < 0 || n_knots < 1 || frames < 0 || (ship0_stride != 0 && ship0_stride < ICEDEM_SHIP_STATE_LEN) ||
        (d_costmap && (rows < 1 || cols < 1 || !(res > 0.0f))) || (d_snap && (snap_every < 1 || n_snap < 1))) {
        set_err(ctx, "icedem_ship_rollouts: bad sizes"); return ICEDEM_E_INVALID;
    }
    if (ctx->hv.p.ship_mode != 1) { set_err(ctx, "icedem_ship_rollouts: context is not in ship_mode 1"); return ICEDEM_E_INVALID; }
    if (S == 0) return ICEDEM_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    const float hl = ctx->hv.p.half_len, hw = ctx->hv.p.half_beam;
    hipLaunchKernelGGL(k_ship_rollouts, dim3((S + 63) / 64), dim3(64), 0, (hipStream_t)stream_, ctx->dv, d_ship0,
                       ship0_stride, d_controls, S, n_knots, hold, frames, d_costmap, rows, cols, x0, y0, res, oob_penalty, hl, hw,
                       d_out, snap_every, n_snap, d_snap);
    HIP_TRY(hipGetLastError());
    return ICEDEM_OK;
}

int icedem_ship_step(IceDemCtx* ctx, float* d_ship, const float* d_setpoint, const float* d_tau_ice,
                     const float* d_v_water_ship, void* stream_) {
    if (!ctx || !d_ship || !d_setpoint || !d_tau_ice) { set_err(ctx, "null argument"); return ICEDEM_E_INVALID; }
    HIP_TRY(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(k_ship_leaf, dim3(1), dim3(64), 0, (hipStream_t)stream_, ctx->dv, d_ship, d_setpoint,
                       d_tau_ice, d_v_water_ship);
    HIP_TRY(hipGetLastError());
    return ICEDEM_OK;
}

}  // extern "C"

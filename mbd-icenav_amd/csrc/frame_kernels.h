// frame_kernels.h -- the per-frame kernels of the DPDEM step.
//
//   k_cull    per element: near-ship test (bounding circles) into a work
//             queue; insertion into the fixed-capacity cell buckets
//   k_near    one wavefront per queued element: hull pass 1 + 2, pre-pass,
//             floe or/max (integer atomics).  The last workgroup to finish
//             (ticket counter) computes the pairwise-tree totals, the thrust
//             scale, floe eligibility and the ship step.
//   k_force   16 lanes per element: activation, submergence, bond forces
//             (lane-parallel terms, sequential CSR-order sum), contacts
//             (lane-parallel scan of the 3x3 buckets, LDS sort by partner),
//             drag, integration
//   k_bond    per bond: fracture; per element: commit + histories; clears
// The same bodies run inside one workgroup per scene for batched rollouts.
#pragma once
#include "kernels.h"

namespace icedem {

constexpr int GRP = 16;                 // lanes per element in k_force
constexpr int EPB = BLOCK / GRP;        // elements per k_force block
constexpr int MAXC2 = 48;               // contact candidates kept per element

// ── broad phase: uniform grid of fixed-capacity buckets ─────────────────────
// Elements of radius <= grid_rsmall sit in the bucket of their cell and scan
// the 3x3 cells around it (cell >= 2 (grid_rsmall + delta)); larger elements
// are inserted into, and scan, every cell within r + grid_rsmall + 2 delta.
// A full bucket sends the element to an overflow list that every element
// also scans.  Any of these gives exactly the reference's contact set.

__device__ __forceinline__ bool is_big(const IceDemParams& P, float r) { return r > P.grid_rsmall; }

__device__ __forceinline__ void cell_box(const IceDemParams& P, float x, float y, float R,
                                         int& x0, int& x1, int& y0, int& y1) {
    float fx0 = floorf((x - R - P.grid_x0) / P.grid_cell), fx1 = floorf((x + R - P.grid_x0) / P.grid_cell);
    float fy0 = floorf((y - R - P.grid_y0) / P.grid_cell), fy1 = floorf((y + R - P.grid_y0) / P.grid_cell);
    const float mx = (float)(P.grid_nx - 1), my = (float)(P.grid_ny - 1);
    x0 = (int)fminf(fmaxf(fx0, 0.0f), mx); x1 = (int)fminf(fmaxf(fx1, 0.0f), mx);
    y0 = (int)fminf(fmaxf(fy0, 0.0f), my); y1 = (int)fminf(fmaxf(fy1, 0.0f), my);
}

__device__ __forceinline__ float big_reach(const IceDemParams& P, float r) {
    return (r + P.grid_rsmall + P.contact_dilation2) * 1.001f + 1e-3f;
}

__device__ __forceinline__ void bucket_put(const DevView* __restrict__ V, int c, int e) {
    const int cap = V->p.grid_bucket;
    const int idx = atomicAdd(&V->cell_count[c], 1);
    if (idx < cap) V->cell_slots[(size_t)c * cap + idx] = e;
    else V->ovf_list[atomicAdd(V->ovf_count, 1)] = e;
}

__device__ __forceinline__ void cells_insert(const DevView* __restrict__ V, int e, float x, float y) {
    const IceDemParams& P = V->p;
    const float r = V->radius[e];
    if (!is_big(P, r)) {
        int cx, cy;
        bucket_put(V, cell_of(P, x, y, cx, cy), e);
    } else {
        int x0, x1, y0, y1;
        cell_box(P, x, y, big_reach(P, r), x0, x1, y0, y1);
        for (int gy = y0; gy <= y1; ++gy)
            for (int gx = x0; gx <= x1; ++gx) bucket_put(V, gy * P.grid_nx + gx, e);
    }
}

// cells to scan for element e: 3x3 around its cell (small) or its box (big)
__device__ __forceinline__ void scan_range(const DevView* __restrict__ V, int e, float x, float y,
                                           int& x0, int& x1, int& y0, int& y1) {
    const IceDemParams& P = V->p;
    const float r = V->radius[e];
    if (!is_big(P, r)) {
        int cx, cy;
        cell_of(P, x, y, cx, cy);
        x0 = max(cx - 1, 0); x1 = min(cx + 1, P.grid_nx - 1);
        y0 = max(cy - 1, 0); y1 = min(cy + 1, P.grid_ny - 1);
    } else {
        cell_box(P, x, y, big_reach(P, r), x0, x1, y0, y1);
    }
}

// One ice-ice pair (lo < hi) of _compute_ice_ice_forces_jax (jax_simulation.py:111-164).
__device__ __forceinline__ bool pair_overlap(const IceDemParams& P, const float* __restrict__ pos,
                                             const float* __restrict__ radius, int n, int lo, int hi) {
    float dx = pos[hi] - pos[lo];
    float dy = pos[n + hi] - pos[n + lo];
    float dist = sqrtf(dx * dx + dy * dy);
    float md = radius[lo] + radius[hi];
    if (P.contact_dilation2 != 0.0f) md = md + P.contact_dilation2;
    float ov = vmax(md - dist, 0.0f);
    ov = (dist > 1e-6f) ? ov : 0.0f;
    return ov > 0.0f;
}

// Contribution of pair (lo, hi) to element `me` (me == lo: i-role, else j-role).
__device__ __forceinline__ void pair_contrib(const IceDemParams& P, const float* __restrict__ pos,
                                             const float* __restrict__ vel, const float* __restrict__ radius,
                                             int n, int lo, int hi, bool me_is_lo,
                                             float& fx, float& fy, float& m) {
    float dx = pos[hi] - pos[lo];
    float dy = pos[n + hi] - pos[n + lo];
    float dist = sqrtf(dx * dx + dy * dy);
    float md = radius[lo] + radius[hi];
    if (P.contact_dilation2 != 0.0f) md = md + P.contact_dilation2;
    float ov = vmax(md - dist, 0.0f);
    ov = (dist > 1e-6f) ? ov : 0.0f;
    float den = vmax(dist, 1e-10f);
    float dirx = dx / den, diry = dy / den;
    float f_n = P.hertz_factor * dm_powf(ov, P.hertz_exponent);
    float fnx = dirx * f_n, fny = diry * f_n;
    float vrx = vel[hi] - vel[lo];
    float vry = vel[n + hi] - vel[n + lo];
    float vn = vrx * dirx + vry * diry;
    float vtx = vrx - vn * dirx, vty = vry - vn * diry;
    float vt = sqrtf(vtx * vtx + vty * vty);
    float vden = vmax(vt, 1e-10f);
    float thx = (-vtx) / vden, thy = (-vty) / vden;
    float f_t = vmin(P.ice_friction * f_n, P.friction_cap);
    float ftx = thx * f_t, fty = thy * f_t;
    if (!(vt > 1e-6f)) { ftx = 0.0f; fty = 0.0f; }
    float fonx = -fnx - ftx, fony = -fny - fty;
    float cross = -(dx * fony - dy * fonx);
    if (me_is_lo) { fx = -fnx + ftx; fy = -fny + fty; m = 0.5f * cross; }
    else { fx = fnx - ftx; fy = fny - fty; m = -0.5f * cross; }
}

// partners after e first (pair_i role, ascending), then partners before e
__device__ __forceinline__ int rot_key(int j, int e, int n) { return j > e ? j - e : j - e + n; }

// ── per-element pre-pass of frame_step (jax_simulation.py:432-479) ───────────

struct PrePass {
    float fm, sx, sy, m;
};

__device__ __forceinline__ PrePass prepass(const IceDemParams& P, float pen, float nxo, float nyo, float r,
                                           float h, float px, float py, float sx, float sy) {
    PrePass o;
    const bool hc = pen > 0.0f;
    float w_c = vmin(P.w_contact_factor * r, P.w_contact_cap);
    float sigma_f = P.flexural_base * sqrtf(h);
    float F_break = sigma_f * w_c * (h * h);
    float pen_crit = P.pen_crit_ref * (h / P.pen_crit_thk_ref);
    float fm = F_break * vmin(pen / vmax(pen_crit, 1e-10f), 1.0f);
    fm = hc ? fm : 0.0f;
    float sxo = fm * nxo;
    float syo = fm * nyo;
    float rel_y = py - sy;
    float side = P.side_push_factor * fm * vsign(rel_y);
    syo = syo + side;
    float rx = px - sx;
    o.fm = fm; o.sx = sxo; o.sy = syo;
    o.m = hc ? (rx * (-syo) - rel_y * (-sxo)) : 0.0f;
    return o;
}


// ── k_cull ───────────────────────────────────────────────────────────────────

// elements [chunk*256, chunk*256+256) by a 256-thread block (all threads call)
__device__ __forceinline__ void cull_body(const DevView* __restrict__ V, const CallDesc* __restrict__ D, int chunk) {
    const IceDemParams& P = V->p;
    const int n = V->n;
    const int e = chunk * BLOCK + threadIdx.x;
    const float* pos = D->s.d_pos;
    const float* ship = D->s.d_ship;
    bool near = false;
    if (e < n) {
        const float px = pos[e], py = pos[n + e];
        const float dx = px - ship[0], dy = py - ship[1];
        const float R = V->bound[e] + P.ship_bound_radius + 1.0f;
        near = (dx * dx + dy * dy) <= R * R;
        V->nearf[e] = near ? 1 : 0;
        if (near) V->near_list[atomicAdd(V->near_count, 1)] = e;
        cells_insert(V, e, px, py);
    }
    const int any = __syncthreads_or(near ? 1 : 0);
    if (threadIdx.x == 0) V->nearblk[chunk] = any ? 1 : 0;
}

__global__ __launch_bounds__(BLOCK) void k_cull(const DevView* __restrict__ V, const CallDesc* __restrict__ D) {
    cull_body(V, D, blockIdx.x);
}

// cell insertion only (sub-steps > 0 and the leaf entry points)
__global__ __launch_bounds__(BLOCK) void k_cellkey(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                                   const float* __restrict__ pos, int from_state) {
    const int n = V->n;
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= n) return;
    if (from_state) pos = D->s.d_pos;
    cells_insert(V, e, pos[e], pos[n + e]);
}

__global__ __launch_bounds__(BLOCK) void k_clear_cells(const DevView* __restrict__ V) {
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t < V->G) V->cell_count[t] = 0;
    if (t == 0) *V->ovf_count = 0;
}

// ── k_near ───────────────────────────────────────────────────────────────────

// waves [wave, wave + nwaves, ...] of the near queue; every lane returns normally
__device__ __forceinline__ void near_body(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                          int wave, int nwaves) {
    const IceDemParams& P = V->p;
    const int n = V->n, maxv = V->maxv;
    const int lane = threadIdx.x & 63;
    const int count = *V->near_count;
    if (wave >= count) return;
    const float* pos = D->s.d_pos;
    const float* ship = D->s.d_ship;
    const float sx = ship[0], sy = ship[1];
    ShipPose SP = make_pose(sx, sy, ship[2], P.half_len, P.half_beam);
    for (int q = wave; q < count; q += nwaves) {
        const int e = V->near_list[q];
        const float px = pos[e], py = pos[n + e];
        float se, ce;
        dm_sincosf(D->s.d_rot[e], &se, &ce);
        float wx = 0.0f, wy = 0.0f;
        bool mv = false;
        if (lane < maxv) {
            float2 lv = V->polys[(size_t)e * maxv + lane];
            wx = ce * lv.x - se * lv.y + px;
            wy = se * lv.x + ce * lv.y + py;
            mv = V->mask[(size_t)e * maxv + lane] != 0;
        }
        float pen, nxo, nyo, cxo, cyo;
        hull_element(SP, P, V->hull, V->S, maxv, wx, wy, mv, pen, nxo, nyo, cxo, cyo);
        if (lane == 0) {
            PrePass o = prepass(P, pen, nxo, nyo, V->radius[e], V->thick[e], px, py, sx, sy);
            V->pen[e] = pen; V->nx[e] = nxo; V->ny[e] = nyo;
            V->fm_raw[e] = o.fm; V->soi_x[e] = o.sx; V->soi_y[e] = o.sy; V->mterm[e] = o.m;
            if (pen > 0.0f) {
                const int f = V->floe[e];
                atomicOr(&V->floe_coll[f], 1u);
                atomicMax(&V->floe_maxfm[f], __float_as_uint(o.fm));
            }
        }
    }
}

// Tree sum of one 256-element block of a near-masked per-element array
// (far elements contribute +0), by one wave: lane l sums its 4 consecutive
// elements pairwise, then the butterfly combines aligned groups.
__device__ __forceinline__ float block256_tree(const DevView* __restrict__ V, const float* __restrict__ a,
                                               int blk, float scale, bool scaled) {
    const int lane = threadIdx.x & 63;
    float v[4];
    #pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int e = blk * BLOCK + lane * 4 + i;
        float x = 0.0f;
        if (e < V->n && V->nearf[e]) x = scaled ? a[e] * scale : a[e];
        v[i] = x;
    }
    float x = (v[0] + v[1]) + (v[2] + v[3]);
    return wave_tree_sum(x);
}

// ship reaction and ship step (jax_simulation.py:471-479, 658-697); thread 0
__device__ __forceinline__ void ship_frame(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                           int k, int last) {
    const IceDemParams& P = V->p;
    float Fx = -V->scal[2], Fy = -V->scal[3];
    float tau[3] = {Fx, Fy, V->scal[1]};
    float* ship = D->s.d_ship;
    float spc[3] = {0.0f, 0.0f, 0.0f}, spn[3] = {0.0f, 0.0f, 0.0f};
    float cu = 0.0f, cr = 0.0f;
    if (P.ship_mode == 1) {
        if (D->in.d_controls && D->in.n_knots > 0) {
            const int knot = min(k / max(D->in.hold, 1), D->in.n_knots - 1);
            cu = D->in.d_controls[2 * knot];
            cr = D->in.d_controls[2 * knot + 1];
        }
    } else {
        const float* sp = D->in.d_setpoints + (size_t)k * 6;
        spc[0] = sp[0]; spc[1] = sp[1]; spc[2] = sp[2];
        spn[0] = sp[3]; spn[1] = sp[4]; spn[2] = sp[5];
        if (P.ice_brake_threshold > 0.0f) {
            float thr = P.ice_brake_threshold;
            float surge = fabsf(tau[0]);
            float bf = vclip(thr / vmax(surge, thr), 0.1f, 1.0f);
            spc[0] = ship[0] + bf * (sp[0] - ship[0]);
            spc[1] = ship[1] + bf * (sp[1] - ship[1]);
        }
    }
    float vwx = 0.0f, vwy = 0.0f;
    if (D->in.d_v_water_ship) { vwx = D->in.d_v_water_ship[2 * k]; vwy = D->in.d_v_water_ship[2 * k + 1]; }
    float st[ICEDEM_SHIP_STATE_LEN];
    for (int i = 0; i < ICEDEM_SHIP_STATE_LEN; ++i) st[i] = ship[i];
    const float Fm = sqrtf(Fx * Fx + Fy * Fy);
    const float J = Fm * P.dt;
    float* ro = D->dg.d_rollout;
    if (ro) {
        if (k == 0) {
            ro[0] = 0.0f; ro[2] = 0.0f;
            ro[3] = 0.5f * P.ship_mass * (st[3] * st[3] + st[4] * st[4]) + 0.5f * P.ship_Izz * (st[5] * st[5]);
        }
        ro[0] = ro[0] + J;
        ro[2] = vmax(ro[2], Fm);
    }
    ship_step(P, st, spc, spn, tau, vwx, vwy, cu, cr);
    for (int i = 0; i < ICEDEM_SHIP_STATE_LEN; ++i) ship[i] = st[i];
    if (ro && last) {
        const float ke = 0.5f * P.ship_mass * (st[3] * st[3] + st[4] * st[4]) + 0.5f * P.ship_Izz * (st[5] * st[5]);
        ro[1] = ro[3] - ke;
    }
    if (D->h.d_eta) { D->h.d_eta[3 * k] = st[0]; D->h.d_eta[3 * k + 1] = st[1]; D->h.d_eta[3 * k + 2] = st[2]; }
    if (D->h.d_J) D->h.d_J[k] = J;
    if (D->h.d_F) { D->h.d_F[2 * k] = Fx; D->h.d_F[2 * k + 1] = Fy; }
    if (last) {
        if (D->dg.d_F_ice_on_ship) { D->dg.d_F_ice_on_ship[0] = Fx; D->dg.d_F_ice_on_ship[1] = Fy; }
        if (D->dg.d_J) D->dg.d_J[0] = J;
    }
}

// Totals, thrust cap (jax_simulation.py:456-479), floe eligibility
// (:511-524) and the ship step, by one workgroup of any size; la, lb hold
// nbpad floats each.
__device__ void mid_body(const DevView* __restrict__ V, const CallDesc* __restrict__ D, float* la, float* lb,
                         int k, int last) {
    const IceDemParams& P = V->p;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int nb = V->nblk, nbpad = V->nbpad;
    for (int b = w; b < nb; b += nw) {
        float a = 0.0f, c = 0.0f, d = 0.0f;
        if (V->nearblk[b]) {
            a = block256_tree(V, V->soi_x, b, 1.0f, false);
            c = block256_tree(V, V->soi_y, b, 1.0f, false);
            d = block256_tree(V, V->mterm, b, 1.0f, false);
        }
        if (lane == 0) { V->part1[b] = a; V->part1[nbpad + b] = c; V->part1[2 * nbpad + b] = d; }
    }
    __syncthreads();
    float tx = lds_tree_sum(V->part1, nb, nbpad, la, lb);
    float ty = lds_tree_sum(V->part1 + nbpad, nb, nbpad, la, lb);
    float tm = lds_tree_sum(V->part1 + 2 * nbpad, nb, nbpad, la, lb);
    float scale = 1.0f;
    if (P.has_thrust_cap) {
        float F_tot = sqrtf(tx * tx + ty * ty);
        scale = (F_tot > P.max_ship_ice_force) ? (P.max_ship_ice_force / vmax(F_tot, 1e-6f)) : 1.0f;
    }
    for (int b = w; b < nb; b += nw) {
        float a = 0.0f, c = 0.0f;
        if (V->nearblk[b]) {
            a = block256_tree(V, V->soi_x, b, scale, P.has_thrust_cap != 0);
            c = block256_tree(V, V->soi_y, b, scale, P.has_thrust_cap != 0);
        }
        if (lane == 0) { V->part2[b] = a; V->part2[nbpad + b] = c; }
    }
    __syncthreads();
    float fx = lds_tree_sum(V->part2, nb, nbpad, la, lb);
    float fy = lds_tree_sum(V->part2 + nbpad, nb, nbpad, la, lb);
    if (threadIdx.x == 0) { V->scal[0] = scale; V->scal[1] = tm; V->scal[2] = fx; V->scal[3] = fy; }
    if (V->nb > 0) {
        uint8_t* elig = D->s.d_floe_elig;
        for (int f = threadIdx.x; f < V->nf; f += blockDim.x) {
            bool coll = V->floe_coll[f] != 0u;
            float mx = __uint_as_float(V->floe_maxfm[f]);
            if (P.has_thrust_cap) mx = mx * scale;
            bool el = (elig[f] != 0) || (coll && (mx > P.elig_threshold));
            elig[f] = el ? 1 : 0;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) ship_frame(V, D, k, last);
    __syncthreads();
}

__global__ __launch_bounds__(BLOCK) void k_near(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                                int k, int last) {
    __shared__ float la[MAX_PARTIALS], lb[MAX_PARTIALS];
    __shared__ int am_last;
    near_body(V, D, blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6), gridDim.x * (BLOCK / 64));
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) am_last = (atomicAdd(V->ticket, 1) == (int)gridDim.x - 1) ? 1 : 0;
    __syncthreads();
    if (!am_last) return;
    __threadfence();
    mid_body(V, D, la, lb, k, last);
    if (threadIdx.x == 0) *V->ticket = 0;
}

// ── k_force ──────────────────────────────────────────────────────────────────

struct ForceLds {
    int cand[EPB][MAXC2];
    int sorted[EPB][MAXC2];
    int keys[EPB][MAXC2];
    float cfx[EPB][MAXC2], cfy[EPB][MAXC2], cfm[EPB][MAXC2];
    int gcell[EPB][GRP];
    int goff[EPB][GRP + 1];
    float bf[EPB][3][GRP];
};

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// group-wide (16 lanes) inclusive prefix sum
__device__ __forceinline__ int group_incl_scan(int v) {
    const int l = threadIdx.x & (GRP - 1);
    #pragma unroll
    for (int off = 1; off < GRP; off <<= 1) {
        int o = __shfl_up(v, off, GRP);
        if (l >= off) v += o;
    }
    return v;
}

// Ice-ice contact force on element e (group-uniform), summed in the
// reference's scatter order; result valid in the group's lane 0.
__device__ void contact_force_group(const DevView* __restrict__ V, const float* __restrict__ pos,
                                    const float* __restrict__ vel, int e, float px, float py,
                                    ForceLds& L, int g, float& fx, float& fy, float& fm) {
    const IceDemParams& P = V->p;
    const int n = V->n;
    const int cap = P.grid_bucket;
    const int l = threadIdx.x & (GRP - 1);
    const int gshift = (threadIdx.x & 63) & ~(GRP - 1);
    int* cand = L.cand[g];
    int x0, x1, y0, y1;
    scan_range(V, e, px, py, x0, x1, y0, y1);
    const int W = x1 - x0 + 1;
    const int ncell = W * (y1 - y0 + 1);
    int cnt = 0;
    auto test = [&](bool have, int j) {
        bool ov = false;
        if (have && j != e) ov = pair_overlap(P, pos, V->radius, n, min(e, j), max(e, j));
        const uint64_t bal = __ballot(ov);
        const uint32_t m = (uint32_t)(bal >> gshift) & ((1u << GRP) - 1u);
        const int before = __builtin_popcount(m & ((1u << l) - 1u));
        if (ov && cnt + before < MAXC2) cand[cnt + before] = j;
        cnt += __builtin_popcount(m);
    };
    for (int c0 = 0; c0 < ncell; c0 += GRP) {
        const int q = c0 + l;
        int cell = 0, cc = 0;
        if (q < ncell) {
            cell = (y0 + q / W) * P.grid_nx + (x0 + q % W);
            cc = min(V->cell_count[cell], cap);
        }
        const int incl = group_incl_scan(cc);
        const int total = __shfl(incl, GRP - 1, GRP);
        L.gcell[g][l] = cell;
        L.goff[g][l] = incl - cc;
        wave_sync_lds();
        for (int base = 0; base < total; base += GRP) {
            const int f = base + l;
            int j = -1;
            if (f < total) {
                int qq = 0;
                #pragma unroll
                for (int t = 1; t < GRP; ++t) qq += (L.goff[g][t] <= f && t < ncell - c0) ? 1 : 0;
                j = V->cell_slots[(size_t)L.gcell[g][qq] * cap + (f - L.goff[g][qq])];
            }
            test(f < total, j);
        }
        wave_sync_lds();
    }
    const int novf = *V->ovf_count;
    for (int base = 0; base < novf; base += GRP) {
        const int f = base + l;
        const int j = (f < novf) ? V->ovf_list[f] : -1;
        test(f < novf, j);
    }
    fx = 0.0f; fy = 0.0f; fm = 0.0f;
    if (cnt <= MAXC2) {
        // sort by rotated partner index; a big element can sit in several of
        // the scanned cells, so duplicates keep their first copy only
        for (int i = l; i < cnt; i += GRP) {
            const int ci = cand[i];
            bool dup = false;
            for (int q = 0; q < i; ++q) dup = dup || (cand[q] == ci);
            L.keys[g][i] = dup ? 0x7fffffff : rot_key(ci, e, n);
        }
        wave_sync_lds();
        int uniq = 0;
        for (int q = 0; q < cnt; ++q) uniq += (L.keys[g][q] != 0x7fffffff) ? 1 : 0;
        for (int i = l; i < cnt; i += GRP) {
            const int ki = L.keys[g][i];
            if (ki == 0x7fffffff) continue;
            int rank = 0;
            for (int q = 0; q < cnt; ++q) rank += (L.keys[g][q] < ki) ? 1 : 0;
            L.sorted[g][rank] = cand[i];
        }
        wave_sync_lds();
        for (int i = l; i < uniq; i += GRP) {
            const int j = L.sorted[g][i];
            float a, b, c;
            pair_contrib(P, pos, vel, V->radius, n, min(e, j), max(e, j), e < j, a, b, c);
            L.cfx[g][i] = a; L.cfy[g][i] = b; L.cfm[g][i] = c;
        }
        wave_sync_lds();
        if (l == 0)
            for (int i = 0; i < uniq; ++i) { fx = fx + L.cfx[g][i]; fy = fy + L.cfy[g][i]; fm = fm + L.cfm[g][i]; }
        wave_sync_lds();
        return;
    }
    // more than MAXC2 candidate hits: exact selection of the next partner by lane 0
    if (l == 0) {
        int last = 0;
        for (;;) {
            int best = 0x7fffffff, bj = -1;
            for (int q = 0; q < ncell; ++q) {
                const int cell = (y0 + q / W) * P.grid_nx + (x0 + q % W);
                const int cc = min(V->cell_count[cell], cap);
                for (int s = 0; s < cc; ++s) {
                    const int j = V->cell_slots[(size_t)cell * cap + s];
                    if (j == e) continue;
                    const int kj = rot_key(j, e, n);
                    if (kj <= last || kj >= best) continue;
                    if (!pair_overlap(P, pos, V->radius, n, min(e, j), max(e, j))) continue;
                    best = kj; bj = j;
                }
            }
            for (int s = 0; s < novf; ++s) {
                const int j = V->ovf_list[s];
                if (j == e) continue;
                const int kj = rot_key(j, e, n);
                if (kj <= last || kj >= best) continue;
                if (!pair_overlap(P, pos, V->radius, n, min(e, j), max(e, j))) continue;
                best = kj; bj = j;
            }
            if (bj < 0) break;
            last = best;
            float a, b, c;
            pair_contrib(P, pos, vel, V->radius, n, min(e, bj), max(e, bj), e < bj, a, b, c);
            fx = fx + a; fy = fy + b; fm = fm + c;
        }
    }
}

// Bond forces on element e (jax_simulation.py:63-108): lanes compute the terms
// of the element's CSR entries, lane 0 adds them in CSR order (the order of
// the reference's two scatter-add passes).  Result valid in lane 0.
__device__ void bond_force_group(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                 const float* __restrict__ pos, int e, int mode, ForceLds& L, int g,
                                 float& fbx, float& fby, float& mb, bool& crack) {
    const IceDemParams& P = V->p;
    const int n = V->n;
    const int l = threadIdx.x & (GRP - 1);
    const int gshift = (threadIdx.x & 63) & ~(GRP - 1);
    const uint8_t* active = D->s.d_active;
    fbx = 0.0f; fby = 0.0f; mb = 0.0f;
    crack = false;
    const int c0 = V->csr_start[e], c1 = V->csr_start[e + 1];
    for (int base = c0; base < c1; base += GRP) {
        const int q = base + l;
        bool broken = false;
        if (q < c1) {
            const int ent = V->csr_ent[q];
            const int b = ent >> 1;
            const bool role_j = (ent & 1) != 0;
            const int i = V->bi[b], j = V->bj[b];
            const bool intact = D->s.d_bond_intact[b] != 0;
            broken = !intact;
            float dx = pos[j] - pos[i];
            float dy = pos[n + j] - pos[n + i];
            float dist = sqrtf(dx * dx + dy * dy);
            float dsafe = vmax(dist, 1e-10f);
            float dirx = dx / dsafe, diry = dy / dsafe;
            float avg_h = (V->thick[i] + V->thick[j]) * 0.5f;
            float stiff = P.youngs_modulus * avg_h * P.bond_stiffness_factor;
            float stretch = dist - V->rest[b];
            float Fm = stiff * stretch * (1.0f - D->s.d_bond_damage[b]);
            float Fx = dirx * Fm, Fy = diry * Fm;
            if (!intact) { Fx = 0.0f; Fy = 0.0f; }
            bool ai, aj;
            if (mode == 1) { ai = active[i] != 0; aj = active[j] != 0; }
            else {
                ai = (active[i] != 0) || (V->floe_coll[V->floe[i]] != 0u);
                aj = (active[j] != 0) || (V->floe_coll[V->floe[j]] != 0u);
            }
            float ok = ((ai && aj) || (!ai && !aj)) ? 1.0f : 0.0f;
            Fx = Fx * ok; Fy = Fy * ok;
            float cross = 0.5f * (dx * Fy - dy * Fx);
            L.bf[g][0][l] = role_j ? -Fx : Fx;
            L.bf[g][1][l] = role_j ? -Fy : Fy;
            L.bf[g][2][l] = role_j ? -cross : cross;
        }
        const uint64_t bal = __ballot(broken);
        crack = crack || (((uint32_t)(bal >> gshift) & ((1u << GRP) - 1u)) != 0u);
        wave_sync_lds();
        if (l == 0) {
            const int m = min(GRP, c1 - base);
            for (int t = 0; t < m; ++t) {
                fbx = fbx + L.bf[g][0][t]; fby = fby + L.bf[g][1][t]; mb = mb + L.bf[g][2][t];
            }
        }
        wave_sync_lds();
    }
}

// elements [chunk*16, chunk*16+16), 16 lanes each (no block barrier inside)
__device__ void force_body(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                           int k, int sub, int src_is_state, int mode,
                           const float* __restrict__ fship_in, float* __restrict__ fout, float* __restrict__ mout,
                           ForceLds& L, int chunk) {
    const IceDemParams& P = V->p;
    const int n = V->n;
    const int g = threadIdx.x / GRP, l = threadIdx.x & (GRP - 1);
    const int e = chunk * EPB + g;
    if (e >= n) return;                       // group-uniform
    const float* pos = src_is_state ? D->s.d_pos : V->nxt_pos;
    const float* vel = src_is_state ? D->s.d_vel : V->nxt_vel;
    const float* rotp = src_is_state ? D->s.d_rot : V->nxt_rot;
    const float* omp = src_is_state ? D->s.d_omega : V->nxt_omega;
    float* dpos = src_is_state ? V->nxt_pos : D->s.d_pos;
    float* dvel = src_is_state ? V->nxt_vel : D->s.d_vel;
    float* drot = src_is_state ? V->nxt_rot : D->s.d_rot;
    float* domg = src_is_state ? V->nxt_omega : D->s.d_omega;
    const uint8_t* active = D->s.d_active;
    const float px = pos[e], py = pos[n + e];
    float fix, fiy, mi;
    contact_force_group(V, pos, vel, e, px, py, L, g, fix, fiy, mi);
    float fbx, fby, mb;
    bool crack;
    bond_force_group(V, D, pos, e, mode, L, g, fbx, fby, mb, crack);
    if (l != 0) return;
    const float h = V->thick[e];
    float shx, shy, subv;
    bool act;
    if (mode == 1) {
        act = active[e] != 0;
        shx = fship_in[e]; shy = fship_in[n + e];
        subv = D->s.d_sub[e];
    } else if (sub == 0) {
        act = (active[e] != 0) || (V->floe_coll[V->floe[e]] != 0u);
        D->s.d_active[e] = act ? 1 : 0;
        const bool nr = V->nearf[e] != 0;
        const float pen = nr ? V->pen[e] : 0.0f;
        const float fmr = nr ? V->fm_raw[e] : 0.0f;
        const float sx_ = nr ? V->soi_x[e] : 0.0f, sy_ = nr ? V->soi_y[e] : 0.0f;
        V->fm[e] = P.has_thrust_cap ? fmr * V->scal[0] : fmr;
        const bool hc = pen > 0.0f;
        const float sub0 = D->s.d_sub[e];
        const float pen_c = vmin(pen, P.submergence_pen_cap);
        float ns = hc ? vmin(sub0 + P.submergence_rate * pen_c * P.dt, h)
                      : vmax(0.0f, sub0 - P.inv_rise_time * sub0 * P.dt);
        D->s.d_sub[e] = ns;
        subv = ns;
        float z = vmin(ns, h);
        float Fb = P.buoyancy_coef * V->area[e] * z;
        shx = sx_ + (hc ? Fb * V->nx[e] : 0.0f);
        shy = sy_ + (hc ? Fb * V->ny[e] : 0.0f);
        if (P.n_ice_steps > 1) { V->shipf[e] = shx; V->shipf[n + e] = shy; }
    } else {
        act = active[e] != 0;
        shx = V->shipf[e]; shy = V->shipf[n + e];
        subv = D->s.d_sub[e];
    }
    if (sub == 0 && mode == 0) V->crack[e] = crack ? 1 : 0;
    // drag (:167-192)
    const float vx = vel[e], vy = vel[n + e];
    float vwx = 0.0f, vwy = 0.0f;
    const float* vwp = D->in.d_v_water;
    if (vwp) {
        const int64_t base = (int64_t)k * D->in.vw_stride_k + (int64_t)e * D->in.vw_stride_e;
        vwx = vwp[base];
        vwy = vwp[base + D->in.vw_stride_c];
    }
    const float area = V->area[e];
    float A_drag = area * (h / P.drag_ref_thickness);
    float vrx = vx - vwx, vry = vy - vwy;
    float vm = sqrtf(vrx * vrx + vry * vry);
    float ms = vmax(vm, 1e-12f);
    float fdx, fdy;
    if (vm > 1e-6f) {
        float qd = P.drag_coef * A_drag * (ms * ms);
        fdx = (vrx / ms) * qd; fdy = (vry / ms) * qd;
    } else {
        float lf = P.drag_coef * A_drag * P.linear_drag_factor;
        fdx = lf * vrx; fdy = lf * vry;
    }
    // rotational drag (:195-212)
    const float r = V->radius[e];
    const float om = omp[e];
    float A_sub = area * vmin(subv / vmax(h, 1e-10f), 1.0f);
    float r3 = r * (r * r);
    float Md = P.rot_drag_coef * r3 * om * fabsf(om) * A_sub;
    const float Fx_t = ((fbx + fix) - fdx) + shx;
    const float Fy_t = ((fby + fiy) - fdy) + shy;
    const float M_t = (mb + mi) + Md;
    if (mode == 1) {
        fout[e] = Fx_t; fout[n + e] = Fy_t; mout[e] = M_t;
        return;
    }
    // integration (:215-268)
    const float dt = P.dt_ice;
    const float vc = P.velocity_clip, pc = P.position_clip;
    const float ro = rotp[e];
    float ax = Fx_t / V->mass[e], ay = Fy_t / V->mass[e];
    float nvx = vclip(vx + ax * dt, -vc, vc);
    float nvy = vclip(vy + ay * dt, -vc, vc);
    float npx = px + nvx * dt;
    float npy = py + nvy * dt;
    float alpha = M_t / V->inertia[e];
    float nw = om * P.angular_damping + alpha * dt;
    nw = vclip(nw, P.ang_cap_lo, P.ang_cap_hi);
    float nr = ro + nw * dt;
    nvx = visfinite(nvx) ? nvx : 0.0f;
    nvy = visfinite(nvy) ? nvy : 0.0f;
    npx = visfinite(npx) ? npx : px;
    npy = visfinite(npy) ? npy : py;
    npx = vclip(npx, -pc, pc);
    npy = vclip(npy, -pc, pc);
    if (!act) { nvx = vx; nvy = vy; npx = px; npy = py; nw = om; nr = ro; }
    dpos[e] = npx; dpos[n + e] = npy;
    dvel[e] = nvx; dvel[n + e] = nvy;
    drot[e] = nr; domg[e] = nw;
}


__global__ __launch_bounds__(BLOCK) void k_force(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                                 int k, int sub, int src_is_state, int mode,
                                                 const float* __restrict__ fship_in,
                                                 float* __restrict__ fout, float* __restrict__ mout) {
    __shared__ ForceLds L;
    force_body(V, D, k, sub, src_is_state, mode, fship_in, fout, mout, L, blockIdx.x);
}

// ── k_bond: fracture (jax_simulation.py:534-656), commit, histories, clears ─

__device__ __forceinline__ void bond_body(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                          int k, int final_in_nxt, int last, int t) {
    const IceDemParams& P = V->p;
    const int n = V->n, nb = V->nb;
    const float* pos = final_in_nxt ? V->nxt_pos : D->s.d_pos;
    const float* vel = final_in_nxt ? V->nxt_vel : D->s.d_vel;
    if (t < nb) {
        const int i = V->bi[t], j = V->bj[t];
        float dx = pos[j] - pos[i];
        float dy = pos[n + j] - pos[n + i];
        float dist = sqrtf(dx * dx + dy * dy);
        float ds_ = vmax(dist, 1e-10f);
        float dirx = dx / ds_, diry = dy / ds_;
        const uint8_t* act = D->s.d_active;
        const uint8_t* elig = D->s.d_floe_elig;
        bool both = act[i] && act[j];
        bool floe_ok = elig[V->floe[i]] && elig[V->floe[j]];
        const float hi_ = V->thick[i], hj = V->thick[j];
        bool breakable = (hi_ <= P.max_fracture_thickness) && (hj <= P.max_fracture_thickness);
        const float fmi = V->fm[i], fmj = V->fm[j];
        bool near = (fmi > 0.0f) || (fmj > 0.0f);
        bool at_crack = V->crack[i] || V->crack[j];
        bool dpdem_ok = both && floe_ok && (near || at_crack) && breakable;
        float vrx = vel[j] - vel[i];
        float vry = vel[n + j] - vel[n + i];
        float vn = vrx * dirx + vry * diry;
        float vtx = vrx - vn * dirx, vty = vry - vn * diry;
        float ds_incr = sqrtf(vtx * vtx + vty * vty) * P.dt_ice;
        bool gate = near && both && floe_ok && breakable;
        const float ds_old = D->s.d_bond_ds[t];
        float new_ds = ds_old + (gate ? ds_incr : 0.0f);
        const float rest = V->rest[t];
        float L0 = vmax(rest, 1e-10f);
        float delta_n = dist - rest;
        float dn_d = near ? delta_n : 0.0f;
        float sn_el = P.k_n * (dn_d / L0);
        float st_el = P.k_s * (new_ds / L0);
        const float dmg = D->s.d_bond_damage[t];
        float sigma_n = sn_el * (1.0f - dmg);
        float sigma_t = st_el * (1.0f - dmg);
        const float dmx = D->s.d_bond_dmmax[t], dm0 = D->s.d_bond_dm0[t];
        const bool intact = D->s.d_bond_intact[t] != 0;
        float nd, ndmx, ndm0;
        bool nint;
        bond_damage_update(delta_n, new_ds, sigma_n, sigma_t, V->sbar[t], P.bond_cohesion, P.bond_friction,
                           P.G_c_I, P.G_c_II, P.eta_bk, dmg, dmx, dm0, intact, nd, ndmx, ndm0, nint);
        float out_d = dpdem_ok ? nd : dmg;
        float out_mx = dpdem_ok ? ndmx : dmx;
        float out_m0 = dpdem_ok ? ndm0 : dm0;
        bool int_d = (dpdem_ok && breakable) ? nint : intact;
        float cf = vmax(fmi, fmj);
        float eff = V->bstr[t] * P.eff_strength_collision;
        bool brk = intact && near && (cf > eff) && breakable;
        bool fin = int_d && !brk;
        D->s.d_bond_damage[t] = out_d;
        D->s.d_bond_dmmax[t] = out_mx;
        D->s.d_bond_dm0[t] = out_m0;
        D->s.d_bond_ds[t] = new_ds;
        D->s.d_bond_intact[t] = fin ? 1 : 0;
        if (D->h.d_bond_intact) D->h.d_bond_intact[(size_t)k * nb + t] = fin ? 1 : 0;
    }
    if (t < n) {
        float px = pos[t], py = pos[n + t];
        if (final_in_nxt) {
            D->s.d_pos[t] = px; D->s.d_pos[n + t] = py;
            D->s.d_vel[t] = V->nxt_vel[t]; D->s.d_vel[n + t] = V->nxt_vel[n + t];
            D->s.d_rot[t] = V->nxt_rot[t]; D->s.d_omega[t] = V->nxt_omega[t];
        }
        if (D->h.d_positions) {
            D->h.d_positions[(size_t)k * 2 * n + t] = px;
            D->h.d_positions[(size_t)k * 2 * n + n + t] = py;
        }
        if (D->h.d_rotations) D->h.d_rotations[(size_t)k * n + t] = final_in_nxt ? V->nxt_rot[t] : D->s.d_rot[t];
        if (D->h.d_submergence) D->h.d_submergence[(size_t)k * n + t] = D->s.d_sub[t];
        if (last) {
            if (D->dg.d_pen_max) D->dg.d_pen_max[t] = V->nearf[t] ? V->pen[t] : 0.0f;
            if (D->dg.d_force_mag) D->dg.d_force_mag[t] = V->fm[t];
        }
    }
    if (t < V->G) V->cell_count[t] = 0;
    if (t < V->nf) { V->floe_coll[t] = 0u; V->floe_maxfm[t] = 0u; }
    if (t == 0) { *V->near_count = 0; *V->ovf_count = 0; }
}

__global__ __launch_bounds__(BLOCK) void k_bond(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                                int k, int final_in_nxt, int last) {
    bond_body(V, D, k, final_in_nxt, last, blockIdx.x * BLOCK + threadIdx.x);
}


// ── batched rollouts: one workgroup simulates one scene for `frames` frames ──

__global__ __launch_bounds__(BLOCK) void k_scene_rollouts(const DevView* __restrict__ views,
                                                          const CallDesc* __restrict__ descs, int frames) {
    __shared__ ForceLds L;
    __shared__ float la[256], lb[256];
    const DevView* V = views + blockIdx.x;
    const CallDesc* D = descs + blockIdx.x;
    const int nblk = V->nblk;
    const int fch = (V->n + EPB - 1) / EPB;
    const int post = max(max(V->nb, V->n), max(V->G, V->nf));
    for (int k = 0; k < frames; ++k) {
        const int last = (k == frames - 1) ? 1 : 0;
        for (int c = 0; c < nblk; ++c) cull_body(V, D, c);
        __syncthreads();
        near_body(V, D, threadIdx.x >> 6, BLOCK / 64);
        __syncthreads();
        mid_body(V, D, la, lb, k, last);
        for (int c = 0; c < fch; ++c)
            force_body(V, D, k, 0, 1, 0, nullptr, nullptr, nullptr, L, c);
        __syncthreads();
        for (int t = threadIdx.x; t < post; t += BLOCK) bond_body(V, D, k, 1, last, t);
        __syncthreads();
    }
}

}  // namespace icedem

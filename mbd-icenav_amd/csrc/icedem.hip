// icedem.hip -- kernels and C ABI of the MI355X DPDEM ice-ship frame step.
// Build: see Makefile (hipcc --offload-arch=gfx950 -O3 -ffp-contract=off).
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>
#include <algorithm>
#include <cmath>

#include "kernels.h"

using namespace icedem;

// ═════════════════════════════════════════════════════════════════════════════
// Kernels
// ═════════════════════════════════════════════════════════════════════════════

// Frame pre-pass: hull narrow phase + contact pre-pass (jax_simulation.py:421-479)
__global__ __launch_bounds__(BLOCK) void k_hull(const DevView* __restrict__ V, const CallDesc* __restrict__ D) {
    __shared__ float red[4];
    const IceDemParams& P = V->p;
    const int n = V->n;
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    const int wave_base = blockIdx.x * BLOCK + (threadIdx.x & ~63);
    const float* pos = D->s.d_pos;
    const float* ship = D->s.d_ship;
    const float sx = ship[0], sy = ship[1], sh = ship[2];
    const bool valid = e < n;
    float px = 0.0f, py = 0.0f, rot = 0.0f;
    if (valid) { px = pos[e]; py = pos[n + e]; rot = D->s.d_rot[e]; }
    float pen, nxo, nyo, cxo, cyo;
    hull_wave(P, V->polys, V->mask, V->bound, V->hull, V->S, V->maxv, wave_base, n, px, py, rot,
              sx, sy, sh, pen, nxo, nyo, cxo, cyo);
    float sxo = 0.0f, syo = 0.0f, mterm = 0.0f;
    if (valid) {
        const float r = V->radius[e], h = V->thick[e];
        const bool hc = pen > 0.0f;
        float w_c = vmin(P.w_contact_factor * r, P.w_contact_cap);
        float sigma_f = P.flexural_base * sqrtf(h);
        float F_break = sigma_f * w_c * (h * h);
        float pen_crit = P.pen_crit_ref * (h / P.pen_crit_thk_ref);
        float fm = F_break * vmin(pen / vmax(pen_crit, 1e-10f), 1.0f);
        fm = hc ? fm : 0.0f;
        sxo = fm * nxo;
        syo = fm * nyo;
        float rel_y = py - sy;
        float side = P.side_push_factor * fm * vsign(rel_y);
        syo = syo + side;
        float rx = px - sx;
        mterm = hc ? (rx * (-syo) - rel_y * (-sxo)) : 0.0f;
        V->pen[e] = pen; V->nx[e] = nxo; V->ny[e] = nyo;
        V->fm_raw[e] = fm; V->soi_x[e] = sxo; V->soi_y[e] = syo;
        if (hc) {
            const int f = V->floe[e];
            atomicOr(&V->floe_coll[f], 1u);
            atomicMax(&V->floe_maxfm[f], __float_as_uint(fm));
        }
        int cx, cy;
        int key = cell_of(P, px, py, cx, cy);
        V->key[e] = key;
        atomicAdd(&V->cell_count[key], 1);
    }
    float a = block_tree_sum<4>(sxo, red);
    float b = block_tree_sum<4>(syo, red);
    float c = block_tree_sum<4>(mterm, red);
    if (threadIdx.x == 0) {
        V->part1[blockIdx.x] = a;
        V->part1[V->nbpad + blockIdx.x] = b;
        V->part1[2 * V->nbpad + blockIdx.x] = c;
    }
}

// Cell histogram only (sub-steps > 0 and the contact-pair leaf).
__global__ __launch_bounds__(BLOCK) void k_cellkey(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                                   const float* __restrict__ pos, int from_state) {
    const int n = V->n;
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= n) return;
    if (from_state) pos = D->s.d_pos;
    int cx, cy;
    int key = cell_of(V->p, pos[e], pos[n + e], cx, cy);
    V->key[e] = key;
    atomicAdd(&V->cell_count[key], 1);
}

// Exclusive scan of V->cell_count into V->cell_start (one workgroup of 1024).
__device__ void scan_cells(const DevView* __restrict__ V, int* lds) {
    const int G = V->G;
    const int T = blockDim.x;
    const int tid = threadIdx.x;
    const int chunk = (G + T - 1) / T;
    const int lo = min(G, tid * chunk), hi = min(G, lo + chunk);
    int s = 0;
    for (int i = lo; i < hi; ++i) s += V->cell_count[i];
    lds[tid] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan over T entries (integers: order-free)
    for (int off = 1; off < T; off <<= 1) {
        int v = (tid >= off) ? lds[tid - off] : 0;
        __syncthreads();
        lds[tid] += v;
        __syncthreads();
    }
    int run = lds[tid] - s;
    for (int i = lo; i < hi; ++i) { V->cell_start[i] = run; run += V->cell_count[i]; }
    if (tid == T - 1) V->cell_start[G] = lds[T - 1];
    __syncthreads();
}

__global__ __launch_bounds__(1024) void k_scan_cells(const DevView* __restrict__ V) {
    __shared__ int lds[1024];
    scan_cells(V, lds);
}

// Thrust cap, floe eligibility and the cell scan (jax_simulation.py:456-524).
__global__ __launch_bounds__(1024) void k_mid(const DevView* __restrict__ V, const CallDesc* __restrict__ D) {
    __shared__ float la[MAX_PARTIALS], lb[MAX_PARTIALS];
    const IceDemParams& P = V->p;
    float tx = lds_tree_sum(V->part1, V->nblk, V->nbpad, la, lb);
    float ty = lds_tree_sum(V->part1 + V->nbpad, V->nblk, V->nbpad, la, lb);
    float tm = lds_tree_sum(V->part1 + 2 * V->nbpad, V->nblk, V->nbpad, la, lb);
    float scale = 1.0f;
    if (P.has_thrust_cap) {
        float F_tot = sqrtf(tx * tx + ty * ty);
        scale = (F_tot > P.max_ship_ice_force) ? (P.max_ship_ice_force / vmax(F_tot, 1e-6f)) : 1.0f;
    }
    if (threadIdx.x == 0) { V->scal[0] = scale; V->scal[1] = tm; }
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
    scan_cells(V, reinterpret_cast<int*>(la));
}

__global__ __launch_bounds__(BLOCK) void k_scatter(const DevView* __restrict__ V) {
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    if (e >= V->n) return;
    const int key = V->key[e];
    const int slot = V->cell_start[key] + atomicAdd(&V->cell_fill[key], 1);
    V->cell_list[slot] = e;
}

// One ice-ice pair (lo < hi) of _compute_ice_ice_forces_jax (jax_simulation.py:111-164).
// Returns false when the pair has no overlap (all contributions exactly zero).
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

__device__ __forceinline__ int rot_key(int j, int e, int n) { return j > e ? j - e : j - e + n; }

// Ice-ice force on element e in the reference scatter order: partners j > e
// ascending (pair_i role), then partners j < e ascending (pair_j role).
__device__ void contact_force(const DevView* __restrict__ V, const float* __restrict__ pos,
                              const float* __restrict__ vel, int e, float px, float py,
                              int* __restrict__ list, float& fx, float& fy, float& fm) {
    const IceDemParams& P = V->p;
    const int n = V->n;
    const int nxg = P.grid_nx, nyg = P.grid_ny;
    int cx, cy;
    cell_of(P, px, py, cx, cy);
    int cnt = 0;
    bool overflow = false;
    for (int gy = max(cy - 1, 0); gy <= min(cy + 1, nyg - 1); ++gy) {
        for (int gx = max(cx - 1, 0); gx <= min(cx + 1, nxg - 1); ++gx) {
            const int c = gy * nxg + gx;
            const int s0 = V->cell_start[c], s1 = V->cell_start[c + 1];
            for (int s = s0; s < s1; ++s) {
                const int j = V->cell_list[s];
                if (j == e) continue;
                const int lo = min(e, j), hi = max(e, j);
                if (!pair_overlap(P, pos, V->radius, n, lo, hi)) continue;
                if (cnt < MAXC) {
                    // insertion by rotated index
                    const int kj = rot_key(j, e, n);
                    int q = cnt;
                    while (q > 0 && rot_key(list[(q - 1) * BLOCK], e, n) > kj) {
                        list[q * BLOCK] = list[(q - 1) * BLOCK];
                        --q;
                    }
                    list[q * BLOCK] = j;
                } else {
                    overflow = true;
                }
                ++cnt;
            }
        }
    }
    fx = 0.0f; fy = 0.0f; fm = 0.0f;
    if (!overflow) {
        for (int q = 0; q < cnt; ++q) {
            const int j = list[q * BLOCK];
            float a, b, c;
            pair_contrib(P, pos, vel, V->radius, n, min(e, j), max(e, j), e < j, a, b, c);
            fx = fx + a; fy = fy + b; fm = fm + c;
        }
        return;
    }
    // slow exact path: repeatedly select the next partner in rotated order
    int last = 0;
    for (int it = 0; it < cnt; ++it) {
        int best = 0x7fffffff, bj = -1;
        for (int gy = max(cy - 1, 0); gy <= min(cy + 1, nyg - 1); ++gy)
            for (int gx = max(cx - 1, 0); gx <= min(cx + 1, nxg - 1); ++gx) {
                const int c = gy * nxg + gx;
                for (int s = V->cell_start[c]; s < V->cell_start[c + 1]; ++s) {
                    const int j = V->cell_list[s];
                    if (j == e) continue;
                    const int kj = rot_key(j, e, n);
                    if (kj <= last || kj >= best) continue;
                    if (!pair_overlap(P, pos, V->radius, n, min(e, j), max(e, j))) continue;
                    best = kj; bj = j;
                }
            }
        if (bj < 0) break;
        last = best;
        float a, b, c;
        pair_contrib(P, pos, vel, V->radius, n, min(e, bj), max(e, bj), e < bj, a, b, c);
        fx = fx + a; fy = fy + b; fm = fm + c;
    }
}

// Sub-step: forces + integration (jax_simulation.py:363-395, 63-268) and, on
// sub-step 0, the per-element part of the pre-pass (:481-509).
// mode 0: integrate into dst; mode 1: write forces/moments to fout/mout.
__global__ __launch_bounds__(BLOCK) void k_force(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                                 int k, int sub, int src_is_state, int mode,
                                                 const float* __restrict__ fship_in,
                                                 float* __restrict__ fout, float* __restrict__ mout) {
    __shared__ int list[MAXC * BLOCK];
    __shared__ float red[4];
    const IceDemParams& P = V->p;
    const int n = V->n;
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    const bool valid = e < n;
    const float* pos = src_is_state ? D->s.d_pos : V->nxt_pos;
    const float* vel = src_is_state ? D->s.d_vel : V->nxt_vel;
    const float* rotp = src_is_state ? D->s.d_rot : V->nxt_rot;
    const float* omp = src_is_state ? D->s.d_omega : V->nxt_omega;
    float* dpos = src_is_state ? V->nxt_pos : D->s.d_pos;
    float* dvel = src_is_state ? V->nxt_vel : D->s.d_vel;
    float* drot = src_is_state ? V->nxt_rot : D->s.d_rot;
    float* domg = src_is_state ? V->nxt_omega : D->s.d_omega;
    const uint8_t* active = D->s.d_active;
    float ssx = 0.0f, ssy = 0.0f;
    if (valid) {
        const float px = pos[e], py = pos[n + e];
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
            const float scale = V->scal[0];
            float fmr = V->fm_raw[e];
            float sx_ = V->soi_x[e], sy_ = V->soi_y[e];
            if (P.has_thrust_cap) {
                V->fm[e] = fmr * scale;
                ssx = sx_ * scale; ssy = sy_ * scale;
            } else {
                V->fm[e] = fmr;
                ssx = sx_; ssy = sy_;
            }
            const float pen = V->pen[e];
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
        // bond forces (jax_simulation.py:63-108) in scatter order
        float fbx = 0.0f, fby = 0.0f, mb = 0.0f;
        bool crack = false;
        const int c0 = V->csr_start[e], c1 = V->csr_start[e + 1];
        for (int q = c0; q < c1; ++q) {
            const int ent = V->csr_ent[q];
            const int b = ent >> 1;
            const bool role_j = (ent & 1) != 0;
            const int i = V->bi[b], j = V->bj[b];
            const bool intact = D->s.d_bond_intact[b] != 0;
            crack = crack || !intact;
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
            if (!role_j) { fbx = fbx + Fx; fby = fby + Fy; mb = mb + cross; }
            else { fbx = fbx + (-Fx); fby = fby + (-Fy); mb = mb + (-cross); }
        }
        if (sub == 0 && mode == 0) V->crack[e] = crack ? 1 : 0;
        // ice-ice contact
        float fix, fiy, mi;
        contact_force(V, pos, vel, e, px, py, &list[threadIdx.x], fix, fiy, mi);
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
            float l = P.drag_coef * A_drag * P.linear_drag_factor;
            fdx = l * vrx; fdy = l * vry;
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
        } else {
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
    }
    if (mode == 0 && sub == 0) {
        float a = block_tree_sum<4>(ssx, red);
        float b = block_tree_sum<4>(ssy, red);
        if (threadIdx.x == 0) {
            V->part2[blockIdx.x] = a;
            V->part2[V->nbpad + blockIdx.x] = b;
        }
    }
}

// Fracture (jax_simulation.py:534-656), commit, histories, clears.
__global__ __launch_bounds__(BLOCK) void k_bond(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                                int k, int final_in_nxt, int last) {
    const IceDemParams& P = V->p;
    const int n = V->n, nb = V->nb;
    const int t = blockIdx.x * BLOCK + threadIdx.x;
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
        bool coll = near;
        float cf = vmax(fmi, fmj);
        float eff = V->bstr[t] * P.eff_strength_collision;
        bool brk = intact && coll && (cf > eff) && breakable;
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
            if (D->dg.d_pen_max) D->dg.d_pen_max[t] = V->pen[t];
            if (D->dg.d_force_mag) D->dg.d_force_mag[t] = V->fm[t];
        }
    }
    if (t < V->G) { V->cell_count[t] = 0; V->cell_fill[t] = 0; }
    if (t < V->nf) { V->floe_coll[t] = 0u; V->floe_maxfm[t] = 0u; }
}

// Ship reaction and ship step (jax_simulation.py:471-479, 658-697).
__global__ __launch_bounds__(64) void k_ship(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                             int k, int last) {
    __shared__ float la[MAX_PARTIALS], lb[MAX_PARTIALS];
    const IceDemParams& P = V->p;
    float sxs = lds_tree_sum(V->part2, V->nblk, V->nbpad, la, lb);
    float sys = lds_tree_sum(V->part2 + V->nbpad, V->nblk, V->nbpad, la, lb);
    if (threadIdx.x != 0) return;
    float Fx = -sxs, Fy = -sys;
    float tau[3] = {Fx, Fy, V->scal[1]};
    float* ship = D->s.d_ship;
    const float* sp = D->in.d_setpoints + (size_t)k * 6;
    float spc[3] = {sp[0], sp[1], sp[2]};
    float spn[3] = {sp[3], sp[4], sp[5]};
    if (P.ice_brake_threshold > 0.0f) {
        float thr = P.ice_brake_threshold;
        float surge = fabsf(tau[0]);
        float bf = vclip(thr / vmax(surge, thr), 0.1f, 1.0f);
        spc[0] = ship[0] + bf * (sp[0] - ship[0]);
        spc[1] = ship[1] + bf * (sp[1] - ship[1]);
    }
    float vwx = 0.0f, vwy = 0.0f;
    if (D->in.d_v_water_ship) { vwx = D->in.d_v_water_ship[2 * k]; vwy = D->in.d_v_water_ship[2 * k + 1]; }
    float st[ICEDEM_SHIP_STATE_LEN];
    for (int i = 0; i < ICEDEM_SHIP_STATE_LEN; ++i) st[i] = ship[i];
    ship_step(P, st, spc, spn, tau, vwx, vwy);
    for (int i = 0; i < ICEDEM_SHIP_STATE_LEN; ++i) ship[i] = st[i];
    float J = sqrtf(Fx * Fx + Fy * Fy) * P.dt;
    if (D->h.d_eta) { D->h.d_eta[3 * k] = st[0]; D->h.d_eta[3 * k + 1] = st[1]; D->h.d_eta[3 * k + 2] = st[2]; }
    if (D->h.d_J) D->h.d_J[k] = J;
    if (D->h.d_F) { D->h.d_F[2 * k] = Fx; D->h.d_F[2 * k + 1] = Fy; }
    if (last) {
        if (D->dg.d_F_ice_on_ship) { D->dg.d_F_ice_on_ship[0] = Fx; D->dg.d_F_ice_on_ship[1] = Fy; }
        if (D->dg.d_J) D->dg.d_J[0] = J;
    }
}

__global__ __launch_bounds__(BLOCK) void k_clear_cells(const DevView* __restrict__ V) {
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t < V->G) { V->cell_count[t] = 0; V->cell_fill[t] = 0; }
}

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

__global__ __launch_bounds__(BLOCK) void k_pair_count(const DevView* __restrict__ V, const float* __restrict__ pos,
                                                      int32_t* cnt) {
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    const int n = V->n;
    if (e >= n) return;
    const IceDemParams& P = V->p;
    int cx, cy;
    cell_of(P, pos[e], pos[n + e], cx, cy);
    int c = 0;
    for (int gy = max(cy - 1, 0); gy <= min(cy + 1, P.grid_ny - 1); ++gy)
        for (int gx = max(cx - 1, 0); gx <= min(cx + 1, P.grid_nx - 1); ++gx) {
            const int cell = gy * P.grid_nx + gx;
            for (int s = V->cell_start[cell]; s < V->cell_start[cell + 1]; ++s) {
                const int j = V->cell_list[s];
                if (j > e && pair_overlap(P, pos, V->radius, n, e, j)) ++c;
            }
        }
    cnt[e] = c;
}

__global__ __launch_bounds__(BLOCK) void k_pair_emit(const DevView* __restrict__ V, const float* __restrict__ pos,
                                                     const int32_t* off, int64_t max_pairs, int32_t* pi, int32_t* pj) {
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    const int n = V->n;
    if (e >= n) return;
    const IceDemParams& P = V->p;
    int cx, cy;
    cell_of(P, pos[e], pos[n + e], cx, cy);
    int o = off[e];
    const int o_end = off[e + 1];
    int last = e;
    // emit partners in ascending order by repeated selection
    while (o < o_end) {
        int best = 0x7fffffff;
        for (int gy = max(cy - 1, 0); gy <= min(cy + 1, P.grid_ny - 1); ++gy)
            for (int gx = max(cx - 1, 0); gx <= min(cx + 1, P.grid_nx - 1); ++gx) {
                const int cell = gy * P.grid_nx + gx;
                for (int s = V->cell_start[cell]; s < V->cell_start[cell + 1]; ++s) {
                    const int j = V->cell_list[s];
                    if (j > last && j < best && pair_overlap(P, pos, V->radius, n, e, j)) best = j;
                }
            }
        if (best == 0x7fffffff) break;
        if (o < max_pairs) { pi[o] = e; pj[o] = best; }
        last = best;
        ++o;
    }
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

// ═════════════════════════════════════════════════════════════════════════════
// Host side
// ═════════════════════════════════════════════════════════════════════════════

struct IceDemCtx {
    int device = 0;
    DevView hv{};              // host copy of the device view
    DevView* dv = nullptr;     // device view
    CallDesc* dd = nullptr;    // device call descriptor
    CallDesc* hd = nullptr;    // pinned host staging for the descriptor
    hipEvent_t desc_done = nullptr;
    std::vector<void*> allocs;
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

template <class T>
static int dalloc(IceDemCtx* ctx, T** p, size_t count) {
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    void* q = nullptr;
    HIP_TRY(hipMalloc(&q, bytes));
    ctx->allocs.push_back(q);
    *p = reinterpret_cast<T*>(q);
    return 0;
}

template <class T>
static int dupload(IceDemCtx* ctx, T** p, const T* src, size_t count) {
    int rc = dalloc(ctx, p, count);
    if (rc) return rc;
    if (count) HIP_TRY(hipMemcpy(*p, src, count * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

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
    HIP_TRY(hipSetDevice(device));
    DevView& H = ctx->hv;
    H.p = *prm;
    H.n = n; H.nb = st->n_bonds; H.nf = st->n_floes; H.maxv = st->max_v; H.S = st->n_hull;
    H.G = prm->grid_nx * prm->grid_ny;
    H.nblk = nblk; H.nbpad = nbpad;
    const size_t nv = (size_t)n * st->max_v;
    int rc = 0;
#define UP(field, type, src, cnt) do { type* _p; rc = dupload(ctx, &_p, (const type*)(src), (cnt)); if (rc) { icedem_destroy(ctx); return rc; } H.field = _p; } while (0)
#define AL(field, type, cnt) do { type* _p; rc = dalloc(ctx, &_p, (cnt)); if (rc) { icedem_destroy(ctx); return rc; } H.field = _p; } while (0)
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
    UP(rest, float, st->h_rest_lengths, st->n_bonds);
    UP(bstr, float, st->h_bond_strengths, st->n_bonds);
    UP(sbar, float, st->h_sigma_bar, st->n_bonds);
    UP(hull, float2, st->h_hull_local, st->n_hull);
    // element -> bond CSR in the reference's scatter order
    {
        std::vector<int32_t> cnt(n + 1, 0), start(n + 1, 0), ent(2 * (size_t)st->n_bonds);
        for (int b = 0; b < st->n_bonds; ++b) { cnt[st->h_bond_i[b]]++; cnt[st->h_bond_j[b]]++; }
        for (int e = 0; e < n; ++e) start[e + 1] = start[e] + cnt[e];
        std::vector<int32_t> fill(start.begin(), start.end() - 1);
        for (int b = 0; b < st->n_bonds; ++b) ent[fill[st->h_bond_i[b]]++] = (b << 1);
        for (int b = 0; b < st->n_bonds; ++b) ent[fill[st->h_bond_j[b]]++] = (b << 1) | 1;
        UP(csr_start, int32_t, start.data(), n + 1);
        UP(csr_ent, int32_t, ent.data(), ent.size());
    }
    AL(pen, float, n); AL(nx, float, n); AL(ny, float, n); AL(fm_raw, float, n);
    AL(soi_x, float, n); AL(soi_y, float, n); AL(fm, float, n); AL(shipf, float, 2 * (size_t)n);
    AL(crack, uint8_t, n); AL(key, int32_t, n);
    AL(cell_count, int32_t, H.G); AL(cell_fill, int32_t, H.G); AL(cell_start, int32_t, H.G + 1);
    AL(cell_list, int32_t, n);
    AL(floe_coll, uint32_t, H.nf); AL(floe_maxfm, uint32_t, H.nf);
    AL(part1, float, 3 * (size_t)nbpad); AL(part2, float, 2 * (size_t)nbpad); AL(scal, float, 4);
    AL(nxt_pos, float, 2 * (size_t)n); AL(nxt_vel, float, 2 * (size_t)n);
    AL(nxt_rot, float, n); AL(nxt_omega, float, n);
    AL(err, int32_t, 4);
#undef UP
#undef AL
    HIP_TRY(hipMemset(H.cell_count, 0, sizeof(int32_t) * H.G));
    HIP_TRY(hipMemset(H.cell_fill, 0, sizeof(int32_t) * H.G));
    HIP_TRY(hipMemset(H.floe_coll, 0, sizeof(uint32_t) * std::max(H.nf, 1)));
    HIP_TRY(hipMemset(H.floe_maxfm, 0, sizeof(uint32_t) * std::max(H.nf, 1)));
    HIP_TRY(hipMemset(H.part1, 0, sizeof(float) * 3 * nbpad));
    HIP_TRY(hipMemset(H.part2, 0, sizeof(float) * 2 * nbpad));
    HIP_TRY(hipMemset(H.err, 0, sizeof(int32_t) * 4));
    {
        DevView* d;
        rc = dalloc(ctx, &d, 1);
        if (rc) { icedem_destroy(ctx); return rc; }
        ctx->dv = d;
        CallDesc* cd;
        rc = dalloc(ctx, &cd, 1);
        if (rc) { icedem_destroy(ctx); return rc; }
        ctx->dd = cd;
    }
    HIP_TRY(hipMemcpy(ctx->dv, &H, sizeof(DevView), hipMemcpyHostToDevice));
    HIP_TRY(hipHostMalloc((void**)&ctx->hd, sizeof(CallDesc), hipHostMallocDefault));
    memset(ctx->hd, 0, sizeof(CallDesc));
    HIP_TRY(hipEventCreateWithFlags(&ctx->desc_done, hipEventDisableTiming));
    HIP_TRY(hipDeviceSynchronize());
    *out = ctx;
    return ICEDEM_OK;
}

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

// ev (optional, profiling): 7 event pairs per frame bracket k_hull, k_mid,
// k_scatter, k_force (all sub-steps with their cell rebuilds), k_bond, k_ship.
static int launch_frames(IceDemCtx* ctx, int K, hipStream_t stream, hipEvent_t* ev = nullptr, float* acc = nullptr) {
    const DevView& H = ctx->hv;
    const DevView* V = ctx->dv;
    const CallDesc* D = ctx->dd;
    const int nblk = H.nblk;
    const int steps = H.p.n_ice_steps;
    const int post = std::max(std::max(H.nb, H.n), std::max(H.G, H.nf));
    const int nbpost = (post + BLOCK - 1) / BLOCK;
    auto mark = [&](int i) { if (ev) (void)hipEventRecord(ev[i], stream); };
    for (int k = 0; k < K; ++k) {
        const int last = (k == K - 1) ? 1 : 0;
        mark(0);
        hipLaunchKernelGGL(k_hull, dim3(nblk), dim3(BLOCK), 0, stream, V, D);
        mark(1);
        hipLaunchKernelGGL(k_mid, dim3(1), dim3(1024), 0, stream, V, D);
        mark(2);
        hipLaunchKernelGGL(k_scatter, dim3(nblk), dim3(BLOCK), 0, stream, V);
        mark(3);
        for (int s = 0; s < steps; ++s) {
            if (s > 0) {
                // rebuild the cell lists for the sub-step's positions
                hipLaunchKernelGGL(k_clear_cells, dim3((H.G + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, stream, V);
                hipLaunchKernelGGL(k_cellkey, dim3(nblk), dim3(BLOCK), 0, stream, V, D,
                                   (const float*)H.nxt_pos, (s % 2 == 0) ? 1 : 0);
                hipLaunchKernelGGL(k_scan_cells, dim3(1), dim3(1024), 0, stream, V);
                hipLaunchKernelGGL(k_scatter, dim3(nblk), dim3(BLOCK), 0, stream, V);
            }
            hipLaunchKernelGGL(k_force, dim3(nblk), dim3(BLOCK), 0, stream, V, D, k, s, (s % 2 == 0) ? 1 : 0, 0,
                               (const float*)nullptr, (float*)nullptr, (float*)nullptr);
        }
        mark(4);
        hipLaunchKernelGGL(k_bond, dim3(nbpost), dim3(BLOCK), 0, stream, V, D, k, (steps % 2 == 1) ? 1 : 0, last);
        mark(5);
        hipLaunchKernelGGL(k_ship, dim3(1), dim3(64), 0, stream, V, D, k, last);
        mark(6);
        if (ev) {
            HIP_TRY(hipEventSynchronize(ev[6]));
            for (int i = 0; i < 6; ++i) {
                float ms = 0.0f;
                HIP_TRY(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
                acc[i] += ms;
            }
        }
    }
    HIP_TRY(hipGetLastError());
    return ICEDEM_OK;
}

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
    if (!use_graph || ctx->hv.p.n_ice_steps > 1) return launch_frames(ctx, K, stream);
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

int icedem_frames_profiled(IceDemCtx* ctx, const IceDemState* state, const IceDemInputs* in, int32_t K,
                           const IceDemHistory* hist, float* h_kernel_ms, void* stream_) {
    if (!ctx || !in || !h_kernel_ms || K < 0) { set_err(ctx, "null argument"); return ICEDEM_E_INVALID; }
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
    hipEvent_t ev[7];
    for (int i = 0; i < 7; ++i) HIP_TRY(hipEventCreate(&ev[i]));
    float acc[6] = {0, 0, 0, 0, 0, 0};
    rc = launch_frames(ctx, K, stream, ev, acc);
    for (int i = 0; i < 7; ++i) (void)hipEventDestroy(ev[i]);
    for (int i = 0; i < 6; ++i) h_kernel_ms[i] = acc[i];
    return rc;
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
    hipLaunchKernelGGL(k_clear_cells, dim3((H.G + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, stream, ctx->dv);
    hipLaunchKernelGGL(k_cellkey, dim3(H.nblk), dim3(BLOCK), 0, stream, ctx->dv, ctx->dd, (const float*)state->d_pos, 0);
    hipLaunchKernelGGL(k_scan_cells, dim3(1), dim3(1024), 0, stream, ctx->dv);
    hipLaunchKernelGGL(k_scatter, dim3(H.nblk), dim3(BLOCK), 0, stream, ctx->dv);
    hipLaunchKernelGGL(k_force, dim3(H.nblk), dim3(BLOCK), 0, stream, ctx->dv, ctx->dd, 0, 0, 1, 1,
                       d_ship_forces, d_forces, d_moments);
    hipLaunchKernelGGL(k_clear_cells, dim3((H.G + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, stream, ctx->dv);
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
    hipLaunchKernelGGL(k_clear_cells, dim3((H.G + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, stream, ctx->dv);
    hipLaunchKernelGGL(k_cellkey, dim3(H.nblk), dim3(BLOCK), 0, stream, ctx->dv, ctx->dd, d_pos, 0);
    hipLaunchKernelGGL(k_scan_cells, dim3(1), dim3(1024), 0, stream, ctx->dv);
    hipLaunchKernelGGL(k_scatter, dim3(H.nblk), dim3(BLOCK), 0, stream, ctx->dv);
    hipLaunchKernelGGL(k_pair_count, dim3(H.nblk), dim3(BLOCK), 0, stream, ctx->dv, d_pos, cnt);
    hipLaunchKernelGGL(k_excl_scan, dim3(1), dim3(1024), 0, stream, H.n, cnt, off);
    hipLaunchKernelGGL(k_pair_emit, dim3(H.nblk), dim3(BLOCK), 0, stream, ctx->dv, d_pos, off, max_pairs,
                       d_pair_i, d_pair_j);
    hipLaunchKernelGGL(k_clear_cells, dim3((H.G + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, stream, ctx->dv);
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

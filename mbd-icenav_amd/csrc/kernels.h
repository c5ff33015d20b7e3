// kernels.h -- DPDEM device functions for gfx950 (wave64) shared by the
// frame kernels (frame_kernels.h) and the leaf entry points: fixed-order
// wave/block reductions, the hull collision (passes 1 and 2 of
// compute_all_collisions_combined), the contact pre-pass, bond, contact,
// drag and integration terms, bond fracture and the ship step.
// No floating-point atomics anywhere: every float sum has a fixed order.
#pragma once
#include "icedem_internal.h"

namespace icedem {

// ── wave / block reductions with a fixed pairwise order ──────────────────────

__device__ __forceinline__ float rl(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// lane l takes the value of the lane DPP control CTRL names
template <int CTRL>
__device__ __forceinline__ int dppi(int x) {
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ float dppf(float x) {
    return __int_as_float(dppi<CTRL>(__float_as_int(x)));
}

// Pairwise tree of the npad (a power of two, at most 64) lane values from
// lane 0 (the tree of tree_sum over npad entries), in every lane.  The
// xor-butterfly pairing: lane i adds lane i ^ 2^k, so after step k every
// lane of an aligned 2^(k+1)-block holds that block's pairwise sum.  Steps 1
// and 2 are DPP quad permutations; for steps 4 and 8 the blocks of the
// previous step hold equal values in every lane, so the half-row and row
// mirrors pair the same sums as xor 4 and xor 8; the rows' sums are combined
// from lanes 0, 16, 32, 48 as the xor-16 and xor-32 steps would ((a + b) and
// (b + a) are the same float).
__device__ __forceinline__ float wave_tree_small(float x, int npad) {
    if (npad > 1) x = x + dppf<0xB1>(x);    // quad_perm [1,0,3,2]: xor 1
    if (npad > 2) x = x + dppf<0x4E>(x);    // quad_perm [2,3,0,1]: xor 2
    if (npad > 4) x = x + dppf<0x141>(x);   // row_half_mirror
    if (npad > 8) x = x + dppf<0x140>(x);   // row_mirror
    if (npad <= 16) return rl(x, 0);
    const float r01 = rl(x, 0) + rl(x, 16);
    if (npad <= 32) return r01;
    return r01 + (rl(x, 32) + rl(x, 48));
}

// pairwise tree of the 64 lane values (tree_sum's order), in every lane
__device__ __forceinline__ float wave_tree_sum(float x) { return wave_tree_small(x, 64); }

// largest of the 64 lane values (order-free), in every lane
__device__ __forceinline__ float wave_max(float x) {
    x = fmaxf(x, dppf<0xB1>(x));
    x = fmaxf(x, dppf<0x4E>(x));
    x = fmaxf(x, dppf<0x141>(x));
    x = fmaxf(x, dppf<0x140>(x));
    return fmaxf(fmaxf(rl(x, 0), rl(x, 16)), fmaxf(rl(x, 32), rl(x, 48)));
}

// Tree sum over the block (blockDim = NW*64); result valid in thread 0.
template <int NW>
__device__ __forceinline__ float block_tree_sum(float x, float* lds) {
    x = wave_tree_sum(x);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) lds[w] = x;
    __syncthreads();
    float r = 0.0f;
    if (threadIdx.x == 0) {
        float a[NW];
        #pragma unroll
        for (int i = 0; i < NW; ++i) a[i] = lds[i];
        #pragma unroll
        for (int len = NW; len > 1; len >>= 1) {
            #pragma unroll
            for (int i = 0; i < len / 2; ++i) a[i] = a[2 * i] + a[2 * i + 1];
        }
        r = a[0];
    }
    __syncthreads();
    return r;
}

// Pairwise tree over src[0..count) zero padded to npad (power of two), done
// by one workgroup with two LDS buffers of npad floats.  Result returned to
// every thread.
__device__ __forceinline__ float lds_tree_sum(const float* src, int count, int npad, float* a, float* b) {
    for (int i = threadIdx.x; i < npad; i += blockDim.x) a[i] = (i < count) ? src[i] : 0.0f;
    __syncthreads();
    int len = npad;
    while (len > 1) {
        int half = len >> 1;
        for (int i = threadIdx.x; i < half; i += blockDim.x) b[i] = a[2 * i] + a[2 * i + 1];
        __syncthreads();
        float* t = a; a = b; b = t;
        len = half;
    }
    float r = a[0];
    __syncthreads();
    return r;
}

// One step of the row-wise argmax: lane l takes lane l - k of its 16-lane row
// (row_shr:k, DPP control ctrl; lanes below k keep their own pair).
template <int CTRL>
__device__ __forceinline__ void argmax_dpp_step(float& val, int& idx) {
    const float ov = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(val), __float_as_int(val), CTRL,
                                                                0xf, 0xf, false));
    const int oi = __builtin_amdgcn_update_dpp(idx, idx, CTRL, 0xf, 0xf, false);
    if (ov > val || (ov == val && oi < idx)) { val = ov; idx = oi; }
}

// First-index argmax across the wave of (val, idx); every lane gets the
// winner.  (value, then smaller index) is a total order on the non-NaN
// pairs, so the order of the reduction does not change the winner: each
// 16-lane row is reduced by DPP row shifts (its last lane holds the row's
// winner), then the four rows' winners are compared in row order.
__device__ __forceinline__ void wave_argmax(float& val, int& idx) {
    argmax_dpp_step<0x111>(val, idx);   // row_shr:1
    argmax_dpp_step<0x112>(val, idx);   // row_shr:2
    argmax_dpp_step<0x114>(val, idx);   // row_shr:4
    argmax_dpp_step<0x118>(val, idx);   // row_shr:8
    float bv = rl(val, 15);
    int bi = __builtin_amdgcn_readlane(idx, 15);
    #pragma unroll
    for (int r = 1; r < 4; ++r) {
        const float v = rl(val, 16 * r + 15);
        const int i = __builtin_amdgcn_readlane(idx, 16 * r + 15);
        if (v > bv || (v == bv && i < bi)) { bv = v; bi = i; }
    }
    val = bv;
    idx = bi;
}

// ── hull narrow phase (jax_geometry.py) ──────────────────────────────────────

struct ShipPose {
    float sx, sy, c_neg, s_neg, c_h, s_h, hl, hb;
};

__device__ __forceinline__ ShipPose make_pose(float sx, float sy, float sh, float hl, float hb) {
    ShipPose P;
    P.sx = sx; P.sy = sy; P.hl = hl; P.hb = hb;
    // sin(-h) = -sin(h) and cos(-h) = cos(h) bit for bit in dm_sincos64:
    // the reduction of -h is exactly the negation of h's (same quadrant
    // count, negated), and the sine series is odd, the cosine series even
    dm_sincosf(sh, &P.s_h, &P.c_h);
    P.s_neg = -P.s_h;
    P.c_neg = P.c_h;
    return P;
}

// point_to_ship_collision_batched (jax_geometry.py:42-134) for one point.
__device__ __forceinline__ void point_to_ship(const ShipPose& P, float px, float py,
                                              float& pen, float& nxo, float& nyo) {
    float dx = px - P.sx, dy = py - P.sy;
    float rel_x = P.c_neg * dx - P.s_neg * dy;
    float rel_y = P.s_neg * dx + P.c_neg * dy;
    float x_rel = rel_x / P.hl;
    bool in_bow = x_rel < -0.6f;
    float t_bow = vclip(((-x_rel) - 0.6f) / 0.4f, 0.0f, 1.0f);
    float lhb = P.hb;
    float nx_bow = 0.0f, ny_bow = 0.0f;
    // the taper only matters where the point can be inside (rel_x <= hl holds
    // in the bow, and lhb <= hb): elsewhere pen, nx, ny are 0 whatever lhb is
    if (in_bow && rel_x >= -P.hl && fabsf(rel_y) <= P.hb) {
        float lhb_bow = P.hb * (1.0f - 0.9f * dm_powf(t_bow, 1.5f));
        lhb = vmax(lhb_bow, P.hb * 0.05f);
        float bow_angle = 0.3f + 0.5f * t_bow;
        float sb, cb;
        dm_sincosf(bow_angle, &sb, &cb);
        nx_bow = -cb;
        ny_bow = vsign(rel_y) * sb;
    }
    bool inside = (rel_x >= -P.hl) && (rel_x <= P.hl) && (fabsf(rel_y) <= lhb);
    float d_bow = rel_x + P.hl;
    float d_side = lhb - fabsf(rel_y);
    float nxb, nyb, pen_raw;
    if (in_bow) {
        nxb = nx_bow; nyb = ny_bow; pen_raw = vmin(d_bow, d_side) + 0.5f;
    } else {
        bool side = d_side <= d_bow;
        nxb = side ? 0.0f : -1.0f;
        nyb = side ? vsign(rel_y) : 0.0f;
        pen_raw = side ? d_side : d_bow;
    }
    float nxw = P.c_h * nxb - P.s_h * nyb;
    float nyw = P.s_h * nxb + P.c_h * nyb;
    pen = inside ? vmax(pen_raw, 0.0f) : 0.0f;
    nxo = inside ? nxw : 0.0f;
    nyo = inside ? nyw : 0.0f;
}

// One element, processed by the whole wave (wave-uniform control flow).
// Lane v holds world vertex v in (wx, wy) and its mask bit in mv.
// Returns pass-1/pass-2 combined (pen, nx, ny, cx, cy) in every lane.
// (ex, ey): the element's position, every world vertex within `reach` of it
// (reach < 0: unknown); pass 2 then tests only the hull points within reach:
// a point farther out lies outside the polygon by more than the vertices'
// rounding, so its ray crossings are even and its penetration 0, and a
// point with penetration 0 cannot win the pass (it counts only above 0).
__device__ __forceinline__ void hull_element(const ShipPose& P, const IceDemParams& prm,
                                             GAS const float2* __restrict__ hull, int S, int maxv,
                                             float wx, float wy, bool mv,
                                             float& pen_o, float& nx_o, float& ny_o,
                                             float& cx_o, float& cy_o,
                                             float ex, float ey, float reach, float2 hl0) {
    const int lane = threadIdx.x & 63;
    const int K = prm.n_edge_samples;
    const int per = 1 + K;
    const int npts = maxv * per;
    // hl0: pass 2's first 64 hull points (hull[min(lane, S - 1)]), loaded
    // by the caller ahead of everything that precedes them here
    // ---- pass 1: vertices + edge samples vs analytic hull (:139-284)
    float bv = -1.0f, bnx = 0.0f, bny = 0.0f, bcx = 0.0f, bcy = 0.0f;
    int bi_ = 0x7fffffff;
    for (int base = 0; base < npts; base += 64) {
        int p = base + lane;
        int kv = p / per;
        int s = p - kv * per;
        int kvc = kv < maxv ? kv : 0;
        int kn = (kvc + 1 == maxv) ? 0 : kvc + 1;
        float vx = __shfl(wx, kvc, 64), vy = __shfl(wy, kvc, 64);
        float ux = __shfl(wx, kn, 64), uy = __shfl(wy, kn, 64);
        int mvi = __shfl((int)mv, kvc, 64);
        if (p < npts) {
            float qx, qy;
            if (s == 0) { qx = vx; qy = vy; }
            else {
                float t = prm.edge_t[s - 1], omt = prm.edge_omt[s - 1];
                qx = vx * omt + ux * t;
                qy = vy * omt + uy * t;
            }
            float pe, nxp, nyp;
            point_to_ship(P, qx, qy, pe, nxp, nyp);
            if (!mvi) pe = 0.0f;
            if (pe > bv) { bv = pe; bi_ = p; bnx = nxp; bny = nyp; bcx = qx; bcy = qy; }
        }
    }
    float v1 = bv; int i1 = bi_;
    wave_argmax(v1, i1);
    int src = i1 & 63;
    float pen1 = v1;
    float nx1 = __shfl(bnx, src, 64), ny1 = __shfl(bny, src, 64);
    float cx1 = __shfl(bcx, src, 64), cy1 = __shfl(bcy, src, 64);

    // ---- pass 2: hull points inside the ice polygon (:289-510)
    float b2v = -1.0f, b2nx = 0.0f, b2ny = 0.0f, b2hx = 0.0f, b2hy = 0.0f;
    int b2i = 0x7fffffff;
    float ccx = 0.0f, ccy = 0.0f;
    bool have_cc = false;
    const float reach2 = reach * reach;
    for (int base = 0; base < S; base += 64) {
        const int s = base + lane;
        const bool inb = s < S;
        const float2 hl = (base == 0) ? hl0 : ldg(hull + (inb ? s : 0));
        const float hx = (hl.x * P.c_h + hl.y * (-P.s_h)) + P.sx;
        const float hy = (hl.x * P.s_h + hl.y * P.c_h) + P.sy;
        bool live = inb;
        if (reach >= 0.0f) {
            const float dx = hx - ex, dy = hy - ey;
            live = live && (dx * dx + dy * dy) <= reach2;
        }
        const uint64_t m = __ballot(live);
        const int nlive = __popcll(m);
        if (nlive == 0) continue;
        if (!have_cc) {
            // masked vertex mean (sequential over the padded vertices)
            float nreal = 0.0f, sxv = 0.0f, syv = 0.0f;
            for (int v = 0; v < maxv; ++v) {
                int mk = __builtin_amdgcn_readlane((int)mv, v);
                float xv = rl(wx, v), yv = rl(wy, v);
                nreal = nreal + (mk ? 1.0f : 0.0f);
                sxv = sxv + (mk ? xv : 0.0f);
                syv = syv + (mk ? yv : 0.0f);
            }
            nreal = vmax(nreal, 1.0f);
            ccx = sxv / nreal; ccy = syv / nreal;
            have_cc = true;
        }
        // GP lanes (a power of two) per live hull point share its polygon
        // edges (kk = sub, sub + GP, ...); crossings add up and the nearest
        // edge is the first minimum, a total order on (distance, edge), so
        // combining the lanes gives the sequential scan's result exactly
        const int GP = nlive > 32 ? 1 : nlive > 16 ? 2 : nlive > 8 ? 4 : nlive > 4 ? 8 : nlive > 2 ? 16 : nlive > 1 ? 32 : 64;
        const int per_round = 64 / GP;
        const int slot = lane / GP, sub = lane & (GP - 1);
        uint64_t rest = m;
        for (int j0 = 0; j0 < nlive; j0 += per_round) {
            // the lane holding live point j0 + slot (points in ascending order)
            int srcl = 0;
            const int cnt = min(per_round, nlive - j0);
            for (int q = 0; q < cnt; ++q) {
                const int bq = __builtin_ctzll(rest);
                rest &= rest - 1ull;
                if (q == slot) srcl = bq;
            }
            const bool have = slot < cnt;
            const float qx = __shfl(hx, srcl, 64), qy = __shfl(hy, srcl, 64);
            int cross = 0;
            float bd = 0.0f;
            int bk = 0x7fffffff;
            for (int k0 = 0; k0 < maxv; k0 += GP) {
                const int kk = k0 + sub;
                const bool ok = kk < maxv;
                const int kkc = ok ? kk : 0;
                int kj = (kkc + 1 == maxv) ? 0 : kkc + 1;
                float xi = __shfl(wx, kkc, 64), yi = __shfl(wy, kkc, 64);
                float xj = __shfl(wx, kj, 64), yj = __shfl(wy, kj, 64);
                bool cy = (yi > qy) != (yj > qy);
                float x_int = (xj - xi) * (qy - yi) / ((yj - yi) + 1e-10f) + xi;
                if (ok && cy && (qx < x_int)) cross++;
                float ex_ = xj - xi, ey_ = yj - yi;
                float wxp = qx - xi, wyp = qy - yi;
                float c1 = wxp * ex_ + wyp * ey_;
                float c2 = (ex_ * ex_ + ey_ * ey_) + 1e-20f;
                float t = vclip(c1 / c2, 0.0f, 1.0f);
                float px_ = xi + t * ex_, py_ = yi + t * ey_;
                float ddx = qx - px_, ddy = qy - py_;
                float d = sqrtf((ddx * ddx + ddy * ddy) + 1e-20f);
                if (ok && (bk == 0x7fffffff || d < bd)) { bd = d; bk = kk; }
            }
            // combine the GP lanes of each point (aligned groups, xor butterfly)
            // (xor 1 to 8 by DPP: after each step an aligned block's lanes
            // hold the same combined value, so the mirrors pair as xor does)
            auto comb = [&](int oc, float od, int ok_) {
                cross += oc;
                if (ok_ != 0x7fffffff && (bk == 0x7fffffff || od < bd || (od == bd && ok_ < bk))) { bd = od; bk = ok_; }
            };
            if (GP > 1) comb(dppi<0xB1>(cross), dppf<0xB1>(bd), dppi<0xB1>(bk));
            if (GP > 2) comb(dppi<0x4E>(cross), dppf<0x4E>(bd), dppi<0x4E>(bk));
            if (GP > 4) comb(dppi<0x141>(cross), dppf<0x141>(bd), dppi<0x141>(bk));
            if (GP > 8) comb(dppi<0x140>(cross), dppf<0x140>(bd), dppi<0x140>(bk));
            for (int off = 16; off < GP; off <<= 1)
                comb(__shfl_xor(cross, off, 64), __shfl_xor(bd, off, 64), __shfl_xor(bk, off, 64));
            if (bk == 0x7fffffff) bk = 0;
            int bkj = (bk + 1 == maxv) ? 0 : bk + 1;
            float xi = __shfl(wx, bk, 64), yi = __shfl(wy, bk, 64);
            float xj = __shfl(wx, bkj, 64), yj = __shfl(wy, bkj, 64);
            if (have && sub == 0) {
                float ex_ = xj - xi, ey_ = yj - yi;
                float nxe = ey_, nye = -ex_;
                float nlen = sqrtf((nxe * nxe + nye * nye) + 1e-20f);
                nxe = nxe / nlen; nye = nye / nlen;
                float midx = (xi + xj) / 2.0f, midy = (yi + yj) / 2.0f;
                float dot = (midx - ccx) * nxe + (midy - ccy) * nye;
                if (dot < 0.0f) { nxe = -nxe; nye = -nye; }
                bool inside = (cross & 1) != 0;
                float ps = inside ? bd : 0.0f;
                float nxs = inside ? nxe : 0.0f, nys = inside ? nye : 0.0f;
                if (ps > b2v) { b2v = ps; b2i = base + srcl; b2nx = nxs; b2ny = nys; b2hx = qx; b2hy = qy; }
            }
        }
    }
    float v2 = b2v; int i2 = b2i;
    wave_argmax(v2, i2);
    int src2 = i2 & 63;
    float pen_i = v2;
    float nx_i = __shfl(b2nx, src2, 64), ny_i = __shfl(b2ny, src2, 64);
    float hx_i = __shfl(b2hx, src2, 64), hy_i = __shfl(b2hy, src2, 64);
    bool upd = pen_i > 0.0f;
    float pen2 = upd ? pen_i : 0.0f;
    float nx2 = upd ? nx_i : 0.0f, ny2 = upd ? ny_i : 0.0f;
    float cx2 = upd ? hx_i : 0.0f, cy2 = upd ? hy_i : 0.0f;
    bool use2 = pen2 > pen1;
    pen_o = use2 ? pen2 : pen1;
    nx_o = use2 ? nx2 : nx1;
    ny_o = use2 ? ny2 : ny1;
    cx_o = use2 ? cx2 : cx1;
    cy_o = use2 ? cy2 : cy1;
}

// Wave-cooperative hull pass over the 64 elements of this wave.  Elements
// farther than bound_e + bound_ship + 1 m from the ship get exact zeros
// (the reference's result for a polygon disjoint from the hull).
__device__ __forceinline__ void hull_wave(const IceDemParams& prm, const float2* __restrict__ polys,
                                          const uint8_t* __restrict__ mask, const float* __restrict__ bound,
                                          const float2* __restrict__ hull, int S, int maxv,
                                          int wave_base, int n, float px, float py, float rot,
                                          float sx, float sy, float sh,
                                          float& pen, float& nxo, float& nyo, float& cxo, float& cyo,
                                          bool want_cp = false) {
    const int lane = threadIdx.x & 63;
    const int e = wave_base + lane;
    pen = 0.0f; nxo = 0.0f; nyo = 0.0f; cxo = 0.0f; cyo = 0.0f;
    bool near = false;
    if (e < n) {
        float dx = px - sx, dy = py - sy;
        float R = bound[e] + prm.ship_bound_radius + 1.0f;
        near = (dx * dx + dy * dy) <= R * R;
        if (want_cp && !near) {
            // no contact: the reference's argmax picks point 0, i.e. the
            // contact point reported is world vertex 0 (jax_geometry.py:276-282)
            float s0, c0;
            dm_sincosf(rot, &s0, &c0);
            float2 v0 = polys[(size_t)e * maxv];
            cxo = c0 * v0.x - s0 * v0.y + px;
            cyo = s0 * v0.x + c0 * v0.y + py;
        }
    }
    uint64_t todo = __ballot(near);
    if (todo == 0) return;
    const float2 hl0 = ldg(G(hull) + (lane < S ? lane : 0));
    ShipPose P = make_pose(sx, sy, sh, prm.half_len, prm.half_beam);
    float se = 0.0f, ce = 1.0f;
    if (near) dm_sincosf(rot, &se, &ce);
    while (todo) {
        int q = __builtin_ctzll(todo);
        todo &= todo - 1;
        int qe = wave_base + q;
        float qpx = rl(px, q), qpy = rl(py, q), qc = rl(ce, q), qs = rl(se, q);
        float wx = 0.0f, wy = 0.0f;
        bool mv = false;
        if (lane < maxv) {
            float2 lv = polys[(size_t)qe * maxv + lane];
            wx = qc * lv.x - qs * lv.y + qpx;
            wy = qs * lv.x + qc * lv.y + qpy;
            mv = mask[(size_t)qe * maxv + lane] != 0;
        }
        float a, b, c, d, f;
        hull_element(P, prm, G(hull), S, maxv, wx, wy, mv, a, b, c, d, f, qpx, qpy, bound[qe] + 0.01f, hl0);
        if (lane == q) { pen = a; nxo = b; nyo = c; cxo = d; cyo = f; }
    }
}

// ── bond damage (jax_bond_fracture.py:29-208) ────────────────────────────────

__device__ __forceinline__ void bond_damage_update(float delta_n, float delta_s, float sigma_n, float sigma_t,
                                                   float sigma_bar, float C, float mu_b, float GcI, float GcII,
                                                   float eta_bk, float damage, float dm_max, float dm_0, bool intact,
                                                   float& nd_o, float& ndmx_o, float& ndm0_o, bool& nint_o) {
    const float SC = 1e12f;
    delta_n = vclip(delta_n, -1e3f, 1e3f);
    delta_s = vclip(delta_s, -1e3f, 1e3f);
    sigma_bar = vclip(sigma_bar, 1e-10f, SC);
    float dn2 = vclip(delta_n, -1e3f, 1e3f);
    float ds2 = vclip(fabsf(delta_s), 0.0f, 1e3f);
    float dnp = vmax(dn2, 0.0f);
    float delta_m = sqrtf(vclip(dnp * dnp + ds2 * ds2 + 1e-20f, 0.0f, 1e12f));
    float new_dm_max = vmax(dm_max, delta_m);
    float tau_bar = vmax(C - mu_b * sigma_n, 1e-6f);
    float sp = vmax(vclip(sigma_n, -SC, SC), 0.0f);
    float tau = vclip(sigma_t, -SC, SC);
    float sb = vmax(vclip(sigma_bar, 1e-10f, SC), 1e-10f);
    float tb = vmax(vclip(tau_bar, 1e-10f, SC), 1e-10f);
    float q1 = sp / sb, q2 = tau / tb;
    float crit = vclip(q1 * q1 + q2 * q2, 0.0f, 1e10f);
    bool init = (crit >= 1.0f) && (dm_0 <= 0.0f) && (delta_m > 1e-5f);
    float new_dm0 = init ? delta_m : dm_0;
    float dnp2 = vmax(delta_n, 0.0f);
    float G_ratio = (delta_s * delta_s) / (dnp2 * dnp2 + 1e-10f);
    float G_I = 1.0f / ((1.0f + G_ratio) + 1e-10f);
    float G_II = G_ratio / ((1.0f + G_ratio) + 1e-10f);
    float G_total = (G_I + G_II) + 1e-10f;
    float ratio = G_II / G_total;
    float G_c = GcI + (GcII - GcI) * dm_powf(ratio, eta_bk);
    float snc = vclip(sigma_n, -SC, SC);
    float stc = vclip(sigma_t, -SC, SC);
    float mp = vmax(snc, 0.0f);
    float sm0_raw = sqrtf(vclip(mp * mp + stc * stc + 1e-20f, 0.0f, 1e24f));
    float sm0 = vmax(sm0_raw, 1e-10f);
    float dmf = (2.0f * G_c) / (fabsf(sm0) + 1e-10f);
    bool meaningful = sm0_raw >= 1e4f;
    bool has_damage = (new_dm0 > 0.0f) && meaningful;
    float d0 = vclip(new_dm0, 0.0f, 1e6f);
    float df = vmax(vclip(dmf, 1e-10f, 1e6f), 1e-10f);
    float dmax = vmax(vclip(new_dm_max, 0.0f, 1e6f), d0 + 1e-10f);
    float denom = dmax * ((df - d0) + 1e-10f);
    denom = vmax(denom, 1e-20f);
    float D = (df * (dmax - d0)) / denom;
    D = vclip(D, 0.0f, 1.0f);
    float computed = has_damage ? D : damage;
    float nd = meaningful ? vmax(computed, damage) : damage;
    nd = vmin(nd, damage + 0.15f);
    nd_o = nd;
    ndmx_o = new_dm_max;
    ndm0_o = new_dm0;
    nint_o = intact && (nd < 1.0f);
}

// ── 3x3 / 3x6 algebra in a fixed order (sequential k accumulation) ──────────

__device__ __forceinline__ void mv3(const float* M, const float* v, float* o) {
    for (int i = 0; i < 3; ++i) {
        float acc = 0.0f;
        for (int k = 0; k < 3; ++k) acc = acc + M[3 * i + k] * v[k];
        o[i] = acc;
    }
}
__device__ __forceinline__ void mm3(const float* A, const float* B, float* o) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            float acc = 0.0f;
            for (int k = 0; k < 3; ++k) acc = acc + A[3 * i + k] * B[3 * k + j];
            o[3 * i + j] = acc;
        }
}

__device__ __forceinline__ float jremf(float x, float y) {
    float r = fmodf(x, y);
    if (r != 0.0f && ((r < 0.0f) != (y < 0.0f))) r = r + y;
    return r;
}

// The part of the ship step that does not depend on the ice forces: DP
// control or direct control, allocation and propeller lag, the current in
// body axes and its damping term (every value ship_post needs).
struct ShipPre {
    float eta[3], nu[3];
    float s, c;                  // sin, cos of the heading
    float tau_act[3];            // actuator forces (body)
    float dn[3];                 // D3 (nu - nu_current)
    float nnew[6], e_new[3];     // propeller states, DP integral
};

__device__ __forceinline__ void ship_pre(const IceDemParams& P, const float* ship, const float* sp, const float* spn,
                                         float vwx, float vwy, float ctrl_u, float ctrl_r, ShipPre& o) {
    const float PI_F = (float)3.141592653589793;        // float32(pi)
    const float TWO_PI_F = (float)6.283185307179586;    // float32(2 pi)
    const float dt = P.dt;
    float* eta = o.eta;
    float* nu = o.nu;
    for (int i = 0; i < 3; ++i) { eta[i] = ship[i]; nu[i] = ship[3 + i]; }
    float nact[6];
    for (int i = 0; i < 6; ++i) nact[i] = ship[6 + i];
    float e_int[3] = {ship[12], ship[13], ship[14]};
    float s, c;
    dm_sincosf(eta[2], &s, &c);
    o.s = s; o.c = c;
    // R^T (row-major)
    float RT[9] = {c, s, 0.0f, -s, c, 0.0f, 0.0f, 0.0f, 1.0f};
    float vw3[3] = {vwx, vwy, 0.0f};
    float nu_c[3];
    mv3(RT, vw3, nu_c);
    float* tau_act = o.tau_act;
    float* nnew = o.nnew;
    float* e_new = o.e_new;
    if (P.ship_mode == 1) {
        const float u = vclip(ctrl_u, -1.0f, 1.0f), rc = vclip(ctrl_r, -1.0f, 1.0f);
        tau_act[0] = u * P.ctrl_thrust_max;
        tau_act[1] = 0.0f;
        tau_act[2] = P.Kd[8] * (rc * P.ctrl_yaw_rate_max - nu[2]);
        for (int i = 0; i < 6; ++i) nnew[i] = nact[i];
        for (int i = 0; i < 3; ++i) e_new[i] = e_int[i];
    } else {
        float dpsi = jremf(spn[2] - sp[2] + PI_F, TWO_PI_F) - PI_F;
        float vd[3] = {(spn[0] - sp[0]) / dt, (spn[1] - sp[1]) / dt, dpsi / dt};
        float nud[3], tff[3];
        mv3(RT, vd, nud);
        mv3(P.D3pKd, nud, tff);
        // dp_control_jax (:127-184)
        float e_psi = jremf(eta[2] - sp[2] + PI_F, TWO_PI_F) - PI_F;
        float e[3] = {eta[0] - sp[0], eta[1] - sp[1], e_psi};
        float RKp[9], RKi[9], a[3], b[3], cc[3];
        mm3(RT, P.Kp, RKp);
        mv3(RKp, e, a);
        mv3(P.Kd, nu, b);
        mm3(RT, P.Ki, RKi);
        mv3(RKi, e_int, cc);
        float tau_c[3];
        for (int i = 0; i < 3; ++i) {
            tau_c[i] = ((-a[i]) - b[i]) - cc[i];
            e_new[i] = e_int[i] + dt * e[i];
        }
        if (P.use_auto_icenav) {
            float t3[3] = {tau_c[0] + tff[0], tau_c[1] + tff[1], tau_c[2] + tff[2]};
            for (int i = 0; i < 6; ++i) {
                float u = 0.0f;
                for (int k = 0; k < 3; ++k) u = u + P.B_pinv[3 * i + k] * t3[k];
                float nc = vsign(u) * sqrtf(fabsf(u) + 1e-12f);
                nc = vclip(nc, -P.n_max[i], P.n_max[i]);
                nnew[i] = nact[i] + dt * (nc - nact[i]) / P.T_n;
            }
            for (int i = 0; i < 3; ++i) {
                float acc = 0.0f;
                for (int k = 0; k < 6; ++k) acc = acc + P.B[6 * i + k] * (fabsf(nnew[k]) * nnew[k]);
                tau_act[i] = acc;
            }
        } else {
            for (int i = 0; i < 3; ++i) tau_act[i] = tau_c[i] + tff[i];
            for (int i = 0; i < 6; ++i) nnew[i] = nact[i];
        }
    }
    // ship_euler_step_jax (:54-84): the damping of the velocity relative to
    // the current
    float nur[3] = {nu[0] - nu_c[0], nu[1] - nu_c[1], nu[2] - nu_c[2]};
    mv3(P.D3, nur, o.dn);
}

// The rest of the ship step once the ice force tau_ice (world frame) is
// known: its body-frame value, the Euler step and the clips.
__device__ __forceinline__ void ship_post(const IceDemParams& P, const ShipPre& o, const float* tau_ice, float* ship) {
    const float PI_F = (float)3.141592653589793;        // float32(pi)
    const float TWO_PI_F = (float)6.283185307179586;    // float32(2 pi)
    const float dt = P.dt;
    const float s = o.s, c = o.c;
    float RT[9] = {c, s, 0.0f, -s, c, 0.0f, 0.0f, 0.0f, 1.0f};
    float tib[3];
    mv3(RT, tau_ice, tib);
    float tot[3] = {o.tau_act[0] + tib[0], o.tau_act[1] + tib[1], o.tau_act[2] + tib[2]};
    float rhs[3], nud2[3];
    for (int i = 0; i < 3; ++i) rhs[i] = tot[i] - o.dn[i];
    mv3(P.M3inv, rhs, nud2);
    float nun[3];
    for (int i = 0; i < 3; ++i) nun[i] = o.nu[i] + dt * nud2[i];
    float vx = c * nun[0] - s * nun[1];
    float vy = s * nun[0] + c * nun[1];
    float psi_new = jremf(o.eta[2] + dt * nun[2] + PI_F, TWO_PI_F) - PI_F;
    ship[0] = o.eta[0] + dt * vx;
    ship[1] = o.eta[1] + dt * vy;
    ship[2] = psi_new;
    ship[3] = vclip(nun[0], -P.vc_surge, P.vc_surge);
    ship[4] = vclip(nun[1], -P.vc_sway, P.vc_sway);
    ship[5] = vclip(nun[2], -P.vc_yaw, P.vc_yaw);
    for (int i = 0; i < 6; ++i) ship[6 + i] = o.nnew[i];
    for (int i = 0; i < 3; ++i) ship[12 + i] = o.e_new[i];
}

// ship_step_jax (jax_vessel_dynamics.py:246-348); ship = eta[3] nu[3] n[6] e_int[3].
// ship_mode 1 (rollouts; no reference counterpart): the thrust is given
// directly, tau = [u X_max, 0, Kd[2,2] (r_cmd r_max - r)] with u, r_cmd in
// [-1, 1]; the Euler step and the velocity clip are the reference's.
__device__ __forceinline__ void ship_step(const IceDemParams& P, float* ship, const float* sp, const float* spn,
                                          const float* tau_ice, float vwx, float vwy,
                                          float ctrl_u = 0.0f, float ctrl_r = 0.0f) {
    ShipPre o;
    ship_pre(P, ship, sp, spn, vwx, vwy, ctrl_u, ctrl_r, o);
    ship_post(P, o, tau_ice, ship);
}

// ── grid cell of a position ──────────────────────────────────────────────────

__device__ __forceinline__ int cell_of(const IceDemParams& P, float x, float y, int& cx, int& cy) {
    float fx = floorf((x - P.grid_x0) / P.grid_cell);
    float fy = floorf((y - P.grid_y0) / P.grid_cell);
    fx = fminf(fmaxf(fx, 0.0f), (float)(P.grid_nx - 1));
    fy = fminf(fmaxf(fy, 0.0f), (float)(P.grid_ny - 1));
    cx = (int)fx; cy = (int)fy;
    return cy * P.grid_nx + cx;
}

}  // namespace icedem

// frame_kernels.h -- the per-frame kernels of the DPDEM step.
//
// A call of K frames (one ice sub-step each) launches
//     k_cull | k_force k_post | k_force k_post | ...
//   k_cull    once per call: near-ship test of every element, the near
//             elements' hull tests inside the culling workgroup, the next
//             frame's near candidates, insertion into the cell buckets of
//             set 0, the partner lists' call-start check.
//   k_force   16 or 8 lanes per element (k_force<GR>): activation,
//             submergence, bond forces, contacts, drag, integration (the
//             integration on one thread per element); frames after a
//             call's first skip the elements frozen through them.  A tail
//             workgroup (dispatched first) forms the pairwise trees of the
//             near elements' ship forces over the blocks holding a near
//             element, the totals, the thrust scale, floe eligibility and the
//             ship step.
//   k_post    per bond: fracture; per element: commit, histories, and the
//             next frame's near test and cell insertion (the bucket set of
//             the next frame's parity); clears the set this frame used; its
//             first workgroups run the next frame's hull tests over the
//             candidates the previous cull listed.
// Frame k uses bucket set k & 1.  Between calls every counter, both bucket
// sets and the floe flags are zero.
// The same bodies run over many contexts in one launch (kb_*) and inside
// one workgroup per scene for batched rollouts.
#pragma once
#include "kernels.h"

namespace icedem {

// k_force gives each element a group of GR lanes, GR = 16 or 8 (a build
// choice per context: 16 shortens one element's chain of dependent loads,
// 8 halves the wave instructions spent per element on its serial parts).
constexpr int GRP = 16;                 // lanes per element of the default force pass
constexpr int MAXC2 = 16;               // contact partners kept per element (more: exact slow path)
template <int GR> constexpr int epb() { return BLOCK / GR; }   // elements per k_force block

__device__ __forceinline__ int32_t* near_count(const DevView* V) { return V->counters; }
// per-floe hull-contact flag and largest contact force of the frames of
// parity `set` (frame k uses k & 1: one frame's near test can set the next
// frame's flags while that frame's k_post clears its own)
__device__ __forceinline__ uint32_t* floe_coll_of(const DevView* V, int set) { return V->floe_coll + (set & 1) * V->nf; }
__device__ __forceinline__ uint32_t* floe_maxfm_of(const DevView* V, int set) { return V->floe_maxfm + (set & 1) * V->nf; }

// ── partner lists: control words ─────────────────────────────────────────────
// Slot k % 3 tells whether a position frame k starts from lies farther than
// sqrt(nl_thr2) from its position at the last list build (NLSUB sub-slots,
// one per XCD, float bits of such a squared displacement, 0 if none).  Frame
// k's force pass reads its slot, its epilogues fill slot k + 1, and frame k's
// k_post reads slots k and k + 1 and clears slot k + 2 for frame k + 1's
// epilogues.  Frame 0 of a call reads the call-start slot instead: the
// call's cull measures the positions the caller passed against the build
// positions (the lists hold every pair that can touch while no element has
// moved farther than that from where its list was built, whoever moved it),
// and marks it when one is farther, when an earlier call's build overflowed,
// or before the context's first build (the slot starts at +inf).
__device__ __forceinline__ uint32_t* nl_slot(const DevView* __restrict__ V, int k) {
    return V->nlctl + (k % 3) * NLSUB;
}
__device__ __forceinline__ uint32_t* nl_ovf(const DevView* __restrict__ V) { return V->nlctl + 3 * NLSUB; }
__device__ __forceinline__ uint32_t* nl_rebuilds(const DevView* __restrict__ V) { return V->nlctl + 3 * NLSUB + 1; }
__device__ __forceinline__ uint32_t* nl_ovf_prev(const DevView* __restrict__ V) { return V->nlctl + 3 * NLSUB + 2; }
__device__ __forceinline__ uint32_t* nl_start(const DevView* __restrict__ V) { return V->nlctl + 4 * NLSUB; }
// frame k rebuilds the lists: some element may have moved far enough that a
// partner outside its list reaches contact.  Plain loads (two 16-byte
// vectors, served by the XCD's L2 after its first wave): a slot is read only
// by kernels launched after the ones that write it.
static_assert(NLSUB == 8, "two uint4 per slot");
__device__ __forceinline__ bool nl_rebuild(const DevView* __restrict__ V, int k) {
    const uint4* s = (const uint4*)(k == 0 ? nl_start(V) : nl_slot(V, k));
    const uint4 a = s[0], b = s[1];
    const uint32_t m = max(max(max(a.x, a.y), max(a.z, a.w)), max(max(b.x, b.y), max(b.z, b.w)));
    return __uint_as_float(m) > V->nl_thr2;
}
// some element had more than NLW candidates at a build of this call: it
// scans the grids every frame, so the buckets are filled every frame (read in
// k_post, after the force pass that may set it)
__device__ __forceinline__ bool nl_overflowed(const DevView* __restrict__ V) { return *nl_ovf(V) != 0u; }

// extra (tail) workgroups of a frame's first k_force: one wave per 256-element block
// element workgroups of k_force for n elements at `lanes` lanes each
__host__ __device__ constexpr int force_chunks(int n, int lanes) { return (n + BLOCK / lanes - 1) / (BLOCK / lanes); }
// lanes per element: the requested 8 or 16, else chosen by element count
__host__ __device__ constexpr int resolve_lanes(int req, int64_t n) {
    return req ? req : (n >= ICEDEM_LANES8_MIN_ELEMENTS ? 8 : 16);
}

// ── broad phase: two uniform grids of fixed-capacity buckets ────────────────
// Elements of radius <= grid_rsmall sit in the bucket of their fine cell
// (cell >= 2 (grid_rsmall + delta)); larger ones sit in the bucket of their
// coarse cell (cell2 >= 2 (2 r_max + delta)).  Every element scans the 2x2
// coarse cells on its side of its own coarse cell for large partners, and
// the fine cells within r + grid_rsmall + delta of its centre for small ones
// (a small partner in contact has its centre that close).  Each element is
// in exactly one bucket, so each partner is seen once.  A full bucket sends the element to
// an overflow list that every element also scans.  Any of these gives
// exactly the reference's contact set.  A slot holds a copy of the element's
// position, velocity and radius, so the contact pass reads its partners from
// the buckets alone.

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

__device__ __forceinline__ void coarse_of(const DevView* __restrict__ V, float x, float y, int& cx, int& cy) {
    const IceDemParams& P = V->p;
    float fx = floorf((x - P.grid_x0) / V->cell2);
    float fy = floorf((y - P.grid_y0) / V->cell2);
    cx = (int)fminf(fmaxf(fx, 0.0f), (float)(V->nx2 - 1));
    cy = (int)fminf(fmaxf(fy, 0.0f), (float)(V->ny2 - 1));
}

// Partial list builds.  A frozen element's list stays exact while every
// element within a list's reach (dreach) of it stays where it was at the
// element's build.  Live elements (the only ones that move within a call)
// stamp the coarse cells they stand in with 1 + the count of builds so far:
// on a build frame (start and end positions), on the frame they go live, and
// whenever they enter another cell than their build position's (staying in
// that cell, it carries the stamp from the build or the activation already).
// The frame before a build marks an element for it when the element is
// active or a cell its reach box touches carries a stamp of the current
// build interval or the one before (a stamp written while the count moves on
// reads either value): any element that can have come within reach since
// the last build stood in one of those cells.  Builds at a call's first
// frame rebuild every list; the call's cull stamps for what the caller may
// have changed between calls (moves, activity).  With every element within
// the threshold of its build position, an element that never stood in a cell
// of a frozen element's box since that element's build stays out of contact.
//
// the stamps' cell of a position: the coarse grid's, by a multiplication
// (stamping and checking use the same monotone map, which is all the
// argument needs)
__device__ __forceinline__ void stamp_xy(const DevView* __restrict__ V, float x, float y, int& cx, int& cy) {
    const IceDemParams& P = V->p;
    const float fx = floorf((x - P.grid_x0) * V->icell2), fy = floorf((y - P.grid_y0) * V->icell2);
    cx = (int)fminf(fmaxf(fx, 0.0f), (float)(V->nx2 - 1));
    cy = (int)fminf(fmaxf(fy, 0.0f), (float)(V->ny2 - 1));
}
__device__ __forceinline__ int coarse_idx(const DevView* __restrict__ V, float x, float y) {
    int cx, cy;
    stamp_xy(V, x, y, cx, cy);
    return cy * V->nx2 + cx;
}
__device__ __forceinline__ bool stamped_near(const DevView* __restrict__ V, float x, float y, uint32_t c) {
    const float R = V->dreach;
    int ax, ay, bx, by;
    stamp_xy(V, x - R, y - R, ax, ay);
    stamp_xy(V, x + R, y + R, bx, by);
    bool hit = !(x == x && y == y);   // NaN position: rebuild
    for (int cy = ay; cy <= by; ++cy)
        for (int cx = ax; cx <= bx; ++cx) hit = hit || G(V->dstamp)[cy * V->nx2 + cx] >= c;
    return hit;
}

// byte offset of bucket record r (fine: r < 2 G) in the records buffer (under
// 4 GiB by construction), the record, and its capacity
__device__ __forceinline__ uint32_t rec_off(const DevView* __restrict__ V, int r) {
    return r < 2 * V->G ? (uint32_t)r * (uint32_t)V->rec_f
                        : (uint32_t)V->coarse_off + (uint32_t)(r - 2 * V->G) * (uint32_t)V->rec_c;
}
__device__ __forceinline__ GAS uint8_t* rec_ptr(const DevView* __restrict__ V, int r) {
    return G(V->cells) + rec_off(V, r);
}
__device__ __forceinline__ int rec_cap(const DevView* __restrict__ V, int r) {
    return r < 2 * V->G ? V->p.grid_bucket : V->cap2;
}
__device__ __forceinline__ int fine_rec(const DevView* __restrict__ V, int set, int cell) { return set * V->G + cell; }
__device__ __forceinline__ int coarse_rec(const DevView* __restrict__ V, int set, int cell) {
    return 2 * V->G + set * V->G2 + cell;
}

struct Slot {
    float x, y, vx, vy, r;
    int e;
};

// a slot is 24 bytes at 8-byte alignment: one 16-byte and one 8-byte access
// (global memory accesses on gfx950 need only 4-byte alignment)
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
typedef float f2u __attribute__((ext_vector_type(2), aligned(4)));

__device__ __forceinline__ Slot slot_load(GAS const uint8_t* __restrict__ rec, int s) {
    GAS const uint8_t* q = rec + 16 + 24 * s;
    const f4u a = *(GAS const f4u*)q;
    const f2u c = *(GAS const f2u*)(q + 16);
    Slot o;
    o.x = a.x; o.y = a.y; o.vx = a.z; o.vy = a.w; o.r = c.x; o.e = __float_as_int(c.y);
    return o;
}

// Bucket record of set `set` for an element of radius r at (x, y).
__device__ __forceinline__ int insert_rec(const DevView* __restrict__ V, int set, float x, float y, float r) {
    const IceDemParams& P = V->p;
    int cx, cy;
    if (!is_big(P, r)) return fine_rec(V, set, cell_of(P, x, y, cx, cy));
    coarse_of(V, x, y, cx, cy);
    return coarse_rec(V, set, cy * V->nx2 + cx);
}

// Element e's slot idx (from the record's insertion count) of record rix:
// the slot, or the overflow list when the bucket is full.
__device__ __forceinline__ void cell_store(const DevView* __restrict__ V, int set, int rix, int idx, int e,
                                           float x, float y, float vx, float vy, float r) {
    if (idx < rec_cap(V, rix)) {
        GAS uint8_t* q = rec_ptr(V, rix) + 16 + 24 * idx;
        f4u a;
        a.x = x; a.y = y; a.z = vx; a.w = vy;
        f2u c;
        c.x = r; c.y = __int_as_float(e);
        *(GAS f4u*)q = a;
        *(GAS f2u*)(q + 16) = c;
    } else {
        G(V->ovf_list)[(size_t)set * V->ovf_cap + atomicAdd(&V->ovf_count[set], 1)] = e;
    }
}

// Element e (state x, y, vx, vy, radius r) into its bucket of set `set`.
__device__ __forceinline__ void cell_insert(const DevView* __restrict__ V, int set, int e,
                                            float x, float y, float vx, float vy, float r) {
    const int rix = insert_rec(V, set, x, y, r);
    cell_store(V, set, rix, atomicAdd((int*)rec_ptr(V, rix), 1), e, x, y, vx, vy, r);
}

// fine cells to scan for an element of radius r at (x, y): those its reach
// box touches (for a small element with cell >= 2 (grid_rsmall + delta) at
// most 3x3, and about 2.5 x 2.5 on average for a median element)
__device__ __forceinline__ void scan_range(const IceDemParams& P, float r, float x, float y,
                                           int& x0, int& x1, int& y0, int& y1) {
    cell_box(P, x, y, big_reach(P, r), x0, x1, y0, y1);
}

// coarse cells to scan: contacts reach at most cell2 / 2, so the 2x2 cells
// on the element's side of its coarse cell hold every large partner
__device__ __forceinline__ void scan_range2(const DevView* __restrict__ V, float x, float y,
                                            int& x0, int& x1, int& y0, int& y1) {
    const IceDemParams& P = V->p;
    const float ux = (x - P.grid_x0) / V->cell2, uy = (y - P.grid_y0) / V->cell2;
    const float fx = floorf(ux), fy = floorf(uy);
    const int cx = (int)fminf(fmaxf(fx, 0.0f), (float)(V->nx2 - 1));
    const int cy = (int)fminf(fmaxf(fy, 0.0f), (float)(V->ny2 - 1));
    const bool lox = (ux - (float)cx) < 0.5f, loy = (uy - (float)cy) < 0.5f;
    x0 = lox ? max(cx - 1, 0) : cx; x1 = lox ? cx : min(cx + 1, V->nx2 - 1);
    y0 = loy ? max(cy - 1, 0) : cy; y1 = loy ? cy : min(cy + 1, V->ny2 - 1);
}

// Overlap test of one ice-ice pair (lo < hi) of _compute_ice_ice_forces_jax
// (jax_simulation.py:111-164), on the pair's values.
__device__ __forceinline__ bool overlap_v(const IceDemParams& P, float xl, float yl, float rlo,
                                          float xh, float yh, float rhi) {
    float dx = xh - xl;
    float dy = yh - yl;
    float dist = sqrtf(dx * dx + dy * dy);
    float md = rlo + rhi;
    if (P.contact_dilation2 != 0.0f) md = md + P.contact_dilation2;
    float ov = vmax(md - dist, 0.0f);
    ov = (dist > 1e-6f) ? ov : 0.0f;
    return ov > 0.0f;
}

__device__ __forceinline__ bool pair_overlap(const IceDemParams& P, GAS const float* __restrict__ pos,
                                             GAS const float* __restrict__ radius, int n, int lo, int hi) {
    return overlap_v(P, pos[lo], pos[n + lo], radius[lo], pos[hi], pos[n + hi], radius[hi]);
}

// Contribution of pair (lo, hi) to one of its elements (me_is_lo: the
// i-role, else the j-role), jax_simulation.py:111-164.
__device__ __forceinline__ void contrib_v(const IceDemParams& P, bool me_is_lo,
                                          float xl, float yl, float vxl, float vyl, float rlo,
                                          float xh, float yh, float vxh, float vyh, float rhi,
                                          float& fx, float& fy, float& m) {
    float dx = xh - xl;
    float dy = yh - yl;
    float dist = sqrtf(dx * dx + dy * dy);
    float md = rlo + rhi;
    if (P.contact_dilation2 != 0.0f) md = md + P.contact_dilation2;
    float ov = vmax(md - dist, 0.0f);
    ov = (dist > 1e-6f) ? ov : 0.0f;
    float den = vmax(dist, 1e-10f);
    float dirx = dx / den, diry = dy / den;
    float f_n = P.hertz_factor * dm_powf(ov, P.hertz_exponent);
    float fnx = dirx * f_n, fny = diry * f_n;
    float vrx = vxh - vxl;
    float vry = vyh - vyl;
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

__device__ __forceinline__ void pair_contrib(const IceDemParams& P, GAS const float* __restrict__ pos,
                                             GAS const float* __restrict__ vel, GAS const float* __restrict__ radius,
                                             int n, int lo, int hi, bool me_is_lo, float& fx, float& fy, float& m) {
    contrib_v(P, me_is_lo, pos[lo], pos[n + lo], vel[lo], vel[n + lo], radius[lo],
              pos[hi], pos[n + hi], vel[hi], vel[n + hi], radius[hi], fx, fy, m);
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

// ── pairwise trees with the order of the oracle's tree_sum ───────────────────

// Tree sum of one 256-element block of a near-masked per-element array
// (far elements contribute +0), by one wave: lane l sums its 4 consecutive
// elements pairwise, then the butterfly combines aligned groups.
__device__ __forceinline__ float block256_tree(const DevView* __restrict__ V, GAS const float* __restrict__ a,
                                               int blk, float scale, bool scaled) {
    const int lane = threadIdx.x & 63;
    float v[4];
    #pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int e = blk * BLOCK + lane * 4 + i;
        float x = 0.0f;
        if (e < V->n) {   // flag and value in one round trip
            const bool nf = G(V->nearf)[e] != 0;
            const float ae = a[e];
            x = nf ? (scaled ? ae * scale : ae) : 0.0f;
        }
        v[i] = x;
    }
    float x = (v[0] + v[1]) + (v[2] + v[3]);
    return wave_tree_sum(x);
}

// Tree sums of one 256-element block of three near-masked per-element arrays
// (the order of block256_tree), every load issued before the first add.
__device__ __forceinline__ void block256_tree3(const DevView* __restrict__ V, GAS const float* __restrict__ a0,
                                               GAS const float* __restrict__ a1, GAS const float* __restrict__ a2,
                                               int blk, float& s0, float& s1, float& s2) {
    const int lane = threadIdx.x & 63;
    const int e0 = blk * BLOCK + lane * 4;
    uint32_t nf = 0;
    float4 x0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), x1 = x0, x2 = x0;
    if (e0 + 3 < V->n) {
        nf = *(GAS const uint32_t*)(G(V->nearf) + e0);
        const f4u y0 = *(GAS const f4u*)(a0 + e0), y1 = *(GAS const f4u*)(a1 + e0), y2 = *(GAS const f4u*)(a2 + e0);
        x0 = make_float4(y0.x, y0.y, y0.z, y0.w);
        x1 = make_float4(y1.x, y1.y, y1.z, y1.w);
        x2 = make_float4(y2.x, y2.y, y2.z, y2.w);
    } else {
        float v0[4] = {0.0f, 0.0f, 0.0f, 0.0f}, v1[4] = {0.0f, 0.0f, 0.0f, 0.0f}, v2[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        #pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (e0 + i < V->n) {
                nf |= (uint32_t)G(V->nearf)[e0 + i] << (8 * i);
                v0[i] = a0[e0 + i]; v1[i] = a1[e0 + i]; v2[i] = a2[e0 + i];
            }
        }
        x0 = make_float4(v0[0], v0[1], v0[2], v0[3]);
        x1 = make_float4(v1[0], v1[1], v1[2], v1[3]);
        x2 = make_float4(v2[0], v2[1], v2[2], v2[3]);
    }
    const bool m0 = (nf & 0xffu) != 0u, m1 = (nf & 0xff00u) != 0u, m2 = (nf & 0xff0000u) != 0u, m3 = (nf >> 24) != 0u;
    auto tree4 = [&](float4 x) {
        return ((m0 ? x.x : 0.0f) + (m1 ? x.y : 0.0f)) + ((m2 ? x.z : 0.0f) + (m3 ? x.w : 0.0f));
    };
    s0 = wave_tree_sum(tree4(x0));
    s1 = wave_tree_sum(tree4(x1));
    s2 = wave_tree_sum(tree4(x2));
}


// ── near test and cell insertion ─────────────────────────────────────────────

// Near-ship test and cell insertion of the elements chunk*256 + threadIdx.x
// for the frame that uses bucket set `set`; (x, y, vx, vy) is the element's
// state at the start of that frame, (sx, sy) the ship's position.  Every
// thread of the 256-thread block calls.
// Near test: the element's centroid within its bound radius + 1 m of the box
// (in the ship's frame) that holds both the analytic hull of pass 1 and the
// hull points of pass 2; elements outside it have no hull contact.  The
// ship's heading comes from a fast sine/cosine: the 1 m margin dwarfs its
// error, and only which elements run the exact hull test depends on it.
// (r, bnd: the element's radius and bound radius, read by the caller with
// its state; queue: list the near elements in the context's near queue, for
// the one-workgroup scene kernel)
// The near test of an element at (x, y) with bound radius bnd against the
// ship at ship[0..2]: its centroid within bnd + margin of the hull's box in
// the ship's frame (margin 1 m: the exact test; larger: a candidate test).
__device__ __forceinline__ bool near_box(const IceDemParams& P, float x, float y, float bnd, GAS const float* ship,
                                         float margin) {
    float sh, ch;
    __sincosf(ship[2], &sh, &ch);
    const float dx = x - ship[0], dy = y - ship[1];
    const float u = ch * dx + sh * dy, v = ch * dy - sh * dx;
    const float b = bnd + margin;
    return fabsf(u) <= P.ship_box_half_len + b && fabsf(v) <= P.ship_box_half_beam + b;
}

// near candidates of the frames of parity p: count counters[1 + p], list
// cand[p n ..], flag cflag[p n + e]
__device__ __forceinline__ int32_t* cand_count(const DevView* V, int p) { return V->counters + 1 + (p & 1); }

// cpar >= 0: also list the elements that pass the near test with margin
// cand_margin as the candidates of the near test one frame later (parity
// cpar); the margin covers one frame's largest displacement of an element
// and of the hull's box, so the candidates hold that test's near elements
// (an element that is near but not listed is caught where it is culled)
__device__ __forceinline__ int cull_chunk(const DevView* __restrict__ V, int chunk, bool valid, int e,
                                          float x, float y, float vx, float vy, float r, float bnd,
                                          GAS const float* ship, int set, bool& near, bool queue,
                                          bool insert = true, bool slot = true, int cpar = -1) {
    const IceDemParams& P = V->p;
    near = false;
    bool cnear = false;
    if (valid) {
        near = near_box(P, x, y, bnd, ship, 1.0f);
        if (cpar >= 0) cnear = near_box(P, x, y, bnd, ship, V->cand_margin);
    }
    if (cpar >= 0) {
        const uint64_t cb = __ballot(cnear);
        if (cb != 0ull) {
            const int lane = threadIdx.x & 63;
            const int cf = __ffsll((unsigned long long)cb) - 1;
            int cbase = 0;
            if (lane == cf) cbase = atomicAdd(cand_count(V, cpar), __popcll(cb));
            cbase = __shfl(cbase, cf, 64);
            if (cnear)
                G(V->cand)[(size_t)(cpar & 1) * V->n + cbase +
                           __builtin_amdgcn_mbcnt_hi((uint32_t)(cb >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)cb, 0u))] = e;
        }
        if (valid) G(V->cflag)[(size_t)(cpar & 1) * V->n + e] = cnear ? 1 : 0;
    }
    // the wave's near elements take consecutive queue entries (one atomic per
    // wave; queue order does not matter), and both atomics are issued before
    // either result is used
    const uint64_t nb = __ballot(near);
    const int lane = threadIdx.x & 63;
    const int first = __ffsll((unsigned long long)nb) - 1;
    int qbase = 0;
    if (queue && lane == first) qbase = atomicAdd(near_count(V), __popcll(nb));
    int rix = 0, idx = 0;
    if (valid && insert) {
        rix = insert_rec(V, set, x, y, r);
        idx = atomicAdd((int*)rec_ptr(V, rix), 1);
    }
    if (queue && nb != 0ull) qbase = __shfl(qbase, first, 64);
    if (valid) {
        G(V->nearf)[e] = near ? 1 : 0;
        if (queue && near) G(V->near_list)[qbase + __builtin_amdgcn_mbcnt_hi((uint32_t)(nb >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)nb, 0u))] = e;
        if (insert) cell_store(V, set, rix, idx, e, x, y, vx, vy, r);
        if (V->nl_on && slot) {   // the element's slot for the partner-list pass
            GAS float* q = G(V->es) + 6 * (size_t)e;
            f4u a;
            a.x = x; a.y = y; a.z = vx; a.w = vy;
            *(GAS f4u*)q = a;
            q[4] = r;
        }
    }
    const int cnt = __syncthreads_count(near ? 1 : 0);
    if (threadIdx.x == 0) G(V->nearcnt)[chunk] = cnt;
    return cnt;
}

// sin/cos of the ship heading for the next near tests (scal[8..11])
__device__ __forceinline__ void store_pose(const DevView* __restrict__ V, const float* __restrict__ ship) {
    const ShipPose SP = make_pose(ship[0], ship[1], ship[2], V->p.half_len, V->p.half_beam);
    G(V->scal)[8] = SP.s_neg; G(V->scal)[9] = SP.c_neg; G(V->scal)[10] = SP.s_h; G(V->scal)[11] = SP.c_h;
}

// cell insertion only (sub-steps > 0 and the leaf entry points)
__global__ __launch_bounds__(BLOCK) void k_cellkey(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                                   const float* __restrict__ pos, const float* __restrict__ vel,
                                                   int from_state, int set) {
    const int n = V->n;
    const int e = blockIdx.x * BLOCK + threadIdx.x;
    const bool valid = e < n;
    if (from_state) { pos = D->s.d_pos; vel = D->s.d_vel; }
    if (valid) cell_insert(V, set, e, pos[e], pos[n + e], vel[e], vel[n + e], G(V->radius)[e]);
}

__device__ __forceinline__ void clear_cells(const DevView* __restrict__ V, int set, int t) {
    if (t < V->G) *(GAS int*)rec_ptr(V, fine_rec(V, set, t)) = 0;
    if (t < V->G2) *(GAS int*)rec_ptr(V, coarse_rec(V, set, t)) = 0;
    if (t == 0) G(V->ovf_count)[set] = 0;
}

__global__ __launch_bounds__(BLOCK) void k_clear_cells(const DevView* __restrict__ V, int set) {
    clear_cells(V, set, blockIdx.x * BLOCK + threadIdx.x);
}

// ── near-ship hull tests ─────────────────────────────────────────────────────

// The static inputs of a near element's hull test (read before the test
// that decides whether it runs, in one round trip with the element's state):
// radius, thickness, floe, reach (every world vertex lies within the
// element's bound of its position, up to the vertices' rounding, which the
// 1 cm margin covers), the lane's body-frame vertex and mask bit, and the
// lane's first hull point.
struct NearIn {
    float rad, thk, reach;
    int fl;
    float2 lv, hl0;
    bool mv;
};

__device__ __forceinline__ NearIn near_in(const DevView* __restrict__ V, int e) {
    const int lane = threadIdx.x & 63, maxv = V->maxv;
    NearIn o;
    o.rad = G(V->radius)[e]; o.thk = G(V->thick)[e];
    o.reach = G(V->bound)[e] + 0.01f;
    o.fl = G(V->floe)[e];
    o.lv = make_float2(0.0f, 0.0f);
    o.mv = false;
    if (lane < maxv) {
        o.lv = ldg(G(V->polys) + (size_t)e * maxv + lane);
        o.mv = G(V->mask)[(size_t)e * maxv + lane] != 0;
    }
    o.hl0 = ldg(G(V->hull) + (lane < V->S ? lane : 0));
    return o;
}

// Hull narrow phase and pre-pass of near element e at position (px, py)
// and rotation rot by the calling wave; lane 0 stores the per-element
// results and the floe or/max of the frames of parity set.
__device__ __forceinline__ void near_elem_at(const DevView* __restrict__ V, const ShipPose& SP, int e, float px,
                                             float py, float rot, int set, const NearIn& ni) {
    const IceDemParams& P = V->p;
    const int maxv = V->maxv;
    const int lane = threadIdx.x & 63;
    float wx = 0.0f, wy = 0.0f;
    float se, ce;
    dm_sincosf(rot, &se, &ce);
    if (lane < maxv) {
        wx = ce * ni.lv.x - se * ni.lv.y + px;
        wy = se * ni.lv.x + ce * ni.lv.y + py;
    }
    float pen, nxo, nyo, cxo, cyo;
    hull_element(SP, P, G(V->hull), V->S, maxv, wx, wy, ni.mv, pen, nxo, nyo, cxo, cyo, px, py, ni.reach, ni.hl0);
    if (lane == 0) {
        PrePass o = prepass(P, pen, nxo, nyo, ni.rad, ni.thk, px, py, SP.sx, SP.sy);
        G(V->pen)[e] = pen; G(V->nx)[e] = nxo; G(V->ny)[e] = nyo;
        G(V->fm_raw)[e] = o.fm;
        G(V->soi_x)[e] = o.sx; G(V->soi_y)[e] = o.sy; G(V->mterm)[e] = o.m;
        if (pen > 0.0f) {
            // results unused: neither atomic is waited for
            (void)atomicOr(floe_coll_of(V, set) + ni.fl, 1u);
            (void)atomicMax(floe_maxfm_of(V, set) + ni.fl, __float_as_uint(o.fm));
        }
    }
}

// near element e with its state read from the call's state buffers
__device__ __forceinline__ void near_elem(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                          const ShipPose& SP, int e, int set) {
    GAS const float* pos = G(D->s.d_pos);
    near_elem_at(V, SP, e, pos[e], pos[V->n + e], G(D->s.d_rot)[e], set, near_in(V, e));
}

// The ship pose of the next near test: position from the ship state, the
// heading's sin/cos as the frame's ship step left them (scal[8..11])
__device__ __forceinline__ ShipPose stored_pose(const DevView* __restrict__ V, const CallDesc* __restrict__ D) {
    const float* ship = D->s.d_ship;
    ShipPose SP;
    SP.sx = ship[0]; SP.sy = ship[1];
    SP.s_neg = G(V->scal)[8]; SP.c_neg = G(V->scal)[9]; SP.s_h = G(V->scal)[10]; SP.c_h = G(V->scal)[11];
    SP.hl = V->p.half_len; SP.hb = V->p.half_beam;
    return SP;
}

// Near tests inside the culling workgroup (a single field's k_cull and
// k_post, where a frame has a handful of near elements): the workgroup's
// near elements are listed in LDS with the position and rotation their
// threads hold, and its waves take them in turn.  Every thread calls, with
// cnt its workgroup's near count (workgroup-uniform).  own_pose: the pose
// from the ship state (a call's first cull, before any ship step stored it).
struct NearLds {
    int cnt;
    int e[BLOCK];
    float x[BLOCK], y[BLOCK], ro[BLOCK];
};

__device__ __forceinline__ void near_inline(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                            NearLds* __restrict__ L, int cnt, bool near, int e, float px, float py,
                                            float ro, int set, bool own_pose) {
    if (cnt == 0) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x == 0) L->cnt = 0;
    __syncthreads();
    const uint64_t nb = __ballot(near);
    if (nb != 0ull) {
        const int first = __ffsll((unsigned long long)nb) - 1;
        int base = 0;
        if (lane == first) base = atomicAdd(&L->cnt, __popcll(nb));
        base = __shfl(base, first, 64);
        if (near) {
            const int q = base + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(nb >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)nb, 0u));
            L->e[q] = e; L->x[q] = px; L->y[q] = py; L->ro[q] = ro;
        }
    }
    __syncthreads();
    ShipPose SP;
    if (own_pose) {
        const float* ship = D->s.d_ship;
        SP = make_pose(ship[0], ship[1], ship[2], V->p.half_len, V->p.half_beam);
    } else {
        SP = stored_pose(V, D);
    }
    for (int q = w; q < cnt; q += BLOCK / 64) {
        const int e = L->e[q];
        near_elem_at(V, SP, e, L->x[q], L->y[q], L->ro[q], set, near_in(V, e));
    }
}


// workgroup lb of a call's first near test (grid: nblk workgroups)
// (with nl: and the near tests of the call's first frame)
__device__ __forceinline__ void cull_kernel_body(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                                 int lb, NearLds* __restrict__ nl) {
    const int n = V->n;
    const int e = lb * BLOCK + threadIdx.x;
    const bool valid = e < n;
    GAS const float* pos = G(D->s.d_pos);
    GAS const float* vel = G(D->s.d_vel);
    float x = 0.0f, y = 0.0f, vx = 0.0f, vy = 0.0f, r = 0.0f, bnd = 0.0f, ro = 0.0f;
    if (valid) {
        x = pos[e]; y = pos[n + e]; vx = vel[e]; vy = vel[n + e];
        r = G(V->radius)[e]; bnd = G(V->bound)[e];
        if (nl) ro = G(D->s.d_rot)[e];   // for the near test, read with the state
    }
    bool near;
    // frame 0 builds the lists when an element lies farther from its build
    // position than the threshold (NaN positions included) or an earlier
    // call's build overflowed a list; the build positions are read with the
    // state
    bool far = false;
    if (V->nl_on && valid) {
        const float bx = G(V->xb)[e], by = G(V->xb)[n + e];
        const float dx = x - bx, dy = y - by;
        far = !(dx * dx + dy * dy <= V->nl_thr2) || *G(nl_ovf_prev(V)) != 0u;
        // what the caller may have changed between calls leaves no stamp
        // of its own: a frozen element away from its build position was
        // moved (a frozen element's build writes where it stands), and an
        // active one may have been activated or moved.  Those stamp the
        // cell they stand in, as an element that moved there in a frame
        // would, so that frozen neighbours whose lists predate the change
        // are rebuilt at the next build (stamped_near).  An element active
        // through the last build stamped its cells then already.
        if (x != bx || y != by || G(D->s.d_active)[e] != 0)
            G(V->dstamp)[coarse_idx(V, x, y)] = *G(nl_rebuilds(V)) + 1u;
    }
    const int cnt = cull_chunk(V, lb, valid, e, x, y, vx, vy, r, bnd, G(D->s.d_ship), 0, near, nl == nullptr, true,
                               true, nl ? 1 : -1);
    if (nl) near_inline(V, D, nl, cnt, near, e, x, y, ro, 0, true);
    if (lb == 0 && threadIdx.x == 0) store_pose(V, D->s.d_ship);
    if (V->nl_on) {
        if (__ballot(far) != 0ull && (threadIdx.x & 63) == 0)
            __hip_atomic_store(nl_start(V) + (lb & (NLSUB - 1)), 0x7f800000u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        // the frame slots and this call's overflow flag start at 0
        if (lb == 0 && threadIdx.x < 3 * NLSUB + 1)
            __hip_atomic_store(V->nlctl + threadIdx.x, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}


// ── ship reaction, floe eligibility, ship step ───────────────────────────────

// ship reaction and ship step (jax_simulation.py:471-479, 658-697); one thread.
// (fsx, fsy) is the sum of the (scaled) ship forces on ice, tm their moment.
// The frame's ship inputs (setpoints, or the rollout's control knot, and the
// current at the ship) for frame k.
__device__ __forceinline__ void ship_inputs(const DevView* __restrict__ V, const CallDesc* __restrict__ D, int k,
                                            float* spc, float* spn, float& cu, float& cr, float& vwx, float& vwy) {
    const IceDemParams& P = V->p;
    for (int i = 0; i < 3; ++i) { spc[i] = 0.0f; spn[i] = 0.0f; }
    cu = 0.0f; cr = 0.0f;
    if (P.ship_mode == 1) {
        if (D->in.d_controls && D->in.n_knots > 0) {
            const int knot = min(k / max(D->in.hold, 1), D->in.n_knots - 1);
            cu = G(D->in.d_controls)[2 * knot];
            cr = G(D->in.d_controls)[2 * knot + 1];
        }
    } else {
        GAS const float* sp = G(D->in.d_setpoints) + (size_t)k * 6;
        spc[0] = sp[0]; spc[1] = sp[1]; spc[2] = sp[2];
        spn[0] = sp[3]; spn[1] = sp[4]; spn[2] = sp[5];
    }
    vwx = 0.0f; vwy = 0.0f;
    if (D->in.d_v_water_ship) { vwx = G(D->in.d_v_water_ship)[2 * k]; vwy = G(D->in.d_v_water_ship)[2 * k + 1]; }
}

// Whether the force-independent part of the ship step can be formed before
// the frame's ice forces are known: not with the ice brake, whose setpoint
// blend depends on them (jax_simulation.py:662-674).
__device__ __forceinline__ bool ship_pre_early(const IceDemParams& P) {
    return !(P.ship_mode != 1 && P.ice_brake_threshold > 0.0f);
}

// ShipPre through device memory, from one kernel to the next
// shpre layout: eta 0-2, nu 3-5, sin 6, cos 7, tau_act 8-10, dn 11-13,
// propeller states 14-19, DP integral 20-22
static_assert(sizeof(ShipPre) == 23 * sizeof(float), "shpre layout covers every field");
template <class P>
__device__ __forceinline__ void ship_pre_get(P g, ShipPre& o) {
    for (int i = 0; i < 3; ++i) { o.eta[i] = g[i]; o.nu[i] = g[3 + i]; o.tau_act[i] = g[8 + i]; o.dn[i] = g[11 + i]; }
    o.s = g[6]; o.c = g[7];
    for (int i = 0; i < 6; ++i) o.nnew[i] = g[14 + i];
    for (int i = 0; i < 3; ++i) o.e_new[i] = g[20 + i];
}

// One workgroup of the kernel before frame k's force pass (a call's first
// cull, or frame k - 1's k_post, after that frame's ship step) forms the
// force-independent part of frame k's ship step, off the force pass's path.
__device__ __forceinline__ void ship_pre_block(const DevView* __restrict__ V, const CallDesc* __restrict__ D, int k) {
    if (threadIdx.x != 0) return;
    float spc[3], spn[3], cu, cr, vwx, vwy;
    ship_inputs(V, D, k, spc, spn, cu, cr, vwx, vwy);
    float st[ICEDEM_SHIP_STATE_LEN];
    GAS const float* ship = G(D->s.d_ship);
    for (int i = 0; i < ICEDEM_SHIP_STATE_LEN; ++i) st[i] = ship[i];
    ShipPre o;
    ship_pre(V->p, st, spc, spn, vwx, vwy, cu, cr, o);
    GAS float* g = G(V->shpre);
    for (int i = 0; i < 3; ++i) { g[i] = o.eta[i]; g[3 + i] = o.nu[i]; g[8 + i] = o.tau_act[i]; g[11 + i] = o.dn[i]; }
    g[6] = o.s; g[7] = o.c;
    for (int i = 0; i < 6; ++i) g[14 + i] = o.nnew[i];
    for (int i = 0; i < 3; ++i) g[20 + i] = o.e_new[i];
}

// shipblk: the grid's last workgroup forms frame 0's ship_pre_block
__global__ __launch_bounds__(BLOCK) void k_cull(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                                int shipblk) {
    __shared__ NearLds nl;
    if (shipblk && (int)blockIdx.x == (int)gridDim.x - 1) {
        ship_pre_block(V, D, 0);
        return;
    }
    cull_kernel_body(V, D, blockIdx.x, &nl);
}

// use_pre: pre holds the force-independent part of this frame's ship step,
// formed earlier (ship_pre_early)
__device__ __forceinline__ void ship_frame(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                           int k, int last, float fsx, float fsy, float tm,
                                           bool use_pre = false, const ShipPre& pre = ShipPre()) {
    const IceDemParams& P = V->p;
    float Fx = -fsx, Fy = -fsy;
    float tau[3] = {Fx, Fy, tm};
    GAS float* ship = G(D->s.d_ship);
    float spc[3] = {0.0f, 0.0f, 0.0f}, spn[3] = {0.0f, 0.0f, 0.0f};
    float cu = 0.0f, cr = 0.0f;
    if (use_pre) {
    } else if (P.ship_mode == 1) {
        if (D->in.d_controls && D->in.n_knots > 0) {
            const int knot = min(k / max(D->in.hold, 1), D->in.n_knots - 1);
            cu = G(D->in.d_controls)[2 * knot];
            cr = G(D->in.d_controls)[2 * knot + 1];
        }
    } else {
        GAS const float* sp = G(D->in.d_setpoints) + (size_t)k * 6;
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
    if (!use_pre && D->in.d_v_water_ship) {
        vwx = G(D->in.d_v_water_ship)[2 * k]; vwy = G(D->in.d_v_water_ship)[2 * k + 1];
    }
    // the ship state (with pre: already in it, as pre.eta and pre.nu, and
    // every entry is rewritten by ship_post)
    float st[ICEDEM_SHIP_STATE_LEN];
    if (!use_pre)
        for (int i = 0; i < ICEDEM_SHIP_STATE_LEN; ++i) st[i] = ship[i];
    const float Fm = sqrtf(Fx * Fx + Fy * Fy);
    const float J = Fm * P.dt;
    GAS float* ro = G(D->dg.d_rollout);
    if (ro) {
        if (k == 0) {
            const float u0 = use_pre ? pre.nu[0] : st[3], v0 = use_pre ? pre.nu[1] : st[4];
            const float r0 = use_pre ? pre.nu[2] : st[5];
            ro[0] = 0.0f; ro[2] = 0.0f;
            ro[3] = 0.5f * P.ship_mass * (u0 * u0 + v0 * v0) + 0.5f * P.ship_Izz * (r0 * r0);
        }
        ro[0] = ro[0] + J;
        ro[2] = vmax(ro[2], Fm);
    }
    if (use_pre) ship_post(P, pre, tau, st);
    else ship_step(P, st, spc, spn, tau, vwx, vwy, cu, cr);
    for (int i = 0; i < ICEDEM_SHIP_STATE_LEN; ++i) ship[i] = st[i];
    if (ro && last) {
        const float ke = 0.5f * P.ship_mass * (st[3] * st[3] + st[4] * st[4]) + 0.5f * P.ship_Izz * (st[5] * st[5]);
        ro[1] = ro[3] - ke;
    }
    if (D->h.d_eta) { G(D->h.d_eta)[3 * k] = st[0]; G(D->h.d_eta)[3 * k + 1] = st[1]; G(D->h.d_eta)[3 * k + 2] = st[2]; }
    if (D->h.d_J) G(D->h.d_J)[k] = J;
    if (D->h.d_F) { G(D->h.d_F)[2 * k] = Fx; G(D->h.d_F)[2 * k + 1] = Fy; }
    if (last) {
        if (D->dg.d_F_ice_on_ship) { G(D->dg.d_F_ice_on_ship)[0] = Fx; G(D->dg.d_F_ice_on_ship)[1] = Fy; }
        if (D->dg.d_J) G(D->dg.d_J)[0] = J;
    }
}

// floe eligibility (jax_simulation.py:511-524); all threads of a workgroup.
// Only floes with a hull contact this frame can change, and the frame's near
// elements with a positive penetration name them (found in the 256-element
// blocks with a near count, a block per wave in turn; a floe named by
// several elements is set as often, to the same value): elig |= max force
// magnitude (thrust-scaled) > threshold.
// Waves w0 .. of the workgroup take part.
__device__ __forceinline__ void floe_eligibility(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                                 float scale, int set, int w0 = 0) {
    const IceDemParams& P = V->p;
    if (V->nb <= 0) return;
    GAS uint8_t* elig = G(D->s.d_floe_elig);
    const int n = V->n, nblk = V->nblk;
    const int lane = threadIdx.x & 63, w = (threadIdx.x >> 6) - w0, nw = (blockDim.x >> 6) - w0;
    for (int base = 64 * w; base < nblk; base += 64 * nw) {
        const int q = base + lane;
        uint64_t m = __ballot(q < nblk && G(V->nearcnt)[q] > 0);
        while (m != 0ull) {
            const int b = base + __builtin_ctzll(m);
            m &= m - 1ull;
            // the block's flags, penetrations and floe ids in one round trip
            uint8_t nf_[BLOCK / 64];
            float pe_[BLOCK / 64];
            int fl_[BLOCK / 64];
            #pragma unroll
            for (int i = 0; i < BLOCK / 64; ++i) {
                const int e = min(b * BLOCK + 64 * i + lane, n - 1);
                nf_[i] = G(V->nearf)[e]; pe_[i] = G(V->pen)[e]; fl_[i] = G(V->floe)[e];
            }
            #pragma unroll
            for (int i = 0; i < BLOCK / 64; ++i) {
                const int e = b * BLOCK + 64 * i + lane;
                if (e < n && nf_[i] != 0 && pe_[i] > 0.0f) {
                    const int f = fl_[i];
                    float mx = __uint_as_float(__hip_atomic_load(floe_maxfm_of(V, set) + f, __ATOMIC_RELAXED,
                                                                 __HIP_MEMORY_SCOPE_AGENT));
                    if (P.has_thrust_cap) mx = mx * scale;
                    if (mx > P.elig_threshold) elig[f] = 1;
                }
            }
        }
    }
}

// Pairwise tree over lv[0..npad) (npad a power of two, at most 64 * 64;
// LDS or memory this workgroup wrote; P the pointer type) by one wave, the
// order of the oracle's tree_sum; result in every lane.  Up to 64 entries:
// one per lane and a butterfly of log2(npad) steps; beyond: each lane reduces
// its contiguous chunk pairwise, up to 16 entries loaded before the first
// add (aligned subtrees of the full tree), then the butterfly.
template <class P>
__device__ __forceinline__ float wg_wave_tree(P lv, int npad) {
    const int lane = threadIdx.x & 63;
    if (npad <= 64) return wave_tree_small(lane < npad ? lv[lane] : 0.0f, npad);
    const int c = npad >> 6;
    const int g = c < 16 ? c : 16;
    P p = lv + (size_t)lane * c;
    float s0 = 0.0f, s1 = 0.0f, r = 0.0f;
    for (int base = 0; base < c; base += 16) {
        float v[16];
        #pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = (i < g) ? p[base + i] : 0.0f;
        #pragma unroll
        for (int st = 1; st < 16; st <<= 1) {
            if (st < g) {
                #pragma unroll
                for (int i = 0; i + st < 16; i += 2 * st) v[i] = v[i] + v[i + st];
            }
        }
        // groups of 16 combined pairwise: (g0 + g1) + (g2 + g3)
        const int q = base >> 4;
        float x = v[0];
        if (q & 1) {
            x = s0 + x;
            if (q & 2) r = s1 + x;
            else s1 = x;
        } else {
            s0 = x;
        }
        r = (c <= 16) ? s0 : (c == 32 ? s1 : r);
    }
    return wave_tree_sum(r);
}

// The tail of a frame's first k_force, one workgroup (dispatched first, so it
// overlaps the element work): the pairwise trees of the near-masked ship
// forces and moment over each 256-element block holding a near element
// (jax_simulation.py:456-470), the other blocks' partials +0, so the totals
// need no hand-off between workgroups; then the thrust scale, the scaled
// partials only when the cap binds (at scale 1 they are the unscaled
// totals, bit for bit), the ship step's final part on wave 0 (preship: its
// force-independent part left by the previous kernel) and the floe
// eligibility on waves 1 .. 3.  The partials (3 nbpad floats) and the list
// of near blocks (nblk ints) live in the workgroup's LDS when they fit
// (`lds`, `cap` bytes of it, unused by the tail otherwise), else in the
// context's partial buffers.
__device__ __forceinline__ void force_tail(const DevView* __restrict__ V, const CallDesc* __restrict__ D, int k,
                                           int last, int set, bool preship, float* lds, int cap) {
    const IceDemParams& P = V->p;
    const int nbpad = V->nbpad, nb = V->nblk;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    constexpr int NW = BLOCK / 64;
    typedef __attribute__((address_space(3))) float lds_f;
    const bool in_lds = 16 * nbpad <= cap;
    float* lv = in_lds ? lds : V->part1;                                 // [3][nbpad] block partials
    int* nlist = in_lds ? (int*)(lds + 3 * nbpad) : (int*)V->part2;      // the near blocks
    // the trees read LDS through LDS pointers (not flat loads)
    auto tree = [&](int a) {
        return in_lds ? wg_wave_tree((lds_f*)(lv + a * nbpad), nbpad) : wg_wave_tree(lv + a * nbpad, nbpad);
    };
    __shared__ int ncnt;
    __shared__ float lpre[32];                    // the ship step's force-independent part
    // The first round trip's loads issue together: the ship step's
    // force-independent part (a value per thread, staged into LDS once the
    // near counts are in) and the near counts of the first NPRE rounds of
    // the listing loop below.
    float pv = 0.0f;
    if (preship && threadIdx.x < 23) pv = G(V->shpre)[threadIdx.x];
    constexpr int NPRE = 4;
    int nc[NPRE];
    #pragma unroll
    for (int i = 0; i < NPRE; ++i) {
        const int q = 64 * (w + NW * i) + lane;
        nc[i] = q < nb ? G(V->nearcnt)[q] : 0;
    }
    if (threadIdx.x == 0) ncnt = 0;
    for (int i = threadIdx.x; i < 3 * nbpad; i += BLOCK) lv[i] = 0.0f;
    __syncthreads();
    // the blocks holding a near element, listed (order free: partials go by index)
    auto list = [&](int base, bool has) {
        const uint64_t m = __ballot(has);
        if (m != 0ull) {
            const int first = __ffsll((unsigned long long)m) - 1;
            int qb = 0;
            if (lane == first) qb = atomicAdd(&ncnt, __popcll(m));
            qb = __shfl(qb, first, 64);
            if (has)
                nlist[qb + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] =
                    base + lane;
        }
    };
    #pragma unroll
    for (int i = 0; i < NPRE; ++i) list(64 * (w + NW * i), nc[i] > 0);
    for (int base = 64 * (w + NW * NPRE); base < nb; base += 64 * NW)
        list(base, base + lane < nb && G(V->nearcnt)[base + lane] > 0);
    if (preship && threadIdx.x < 23) lpre[threadIdx.x] = pv;
    __syncthreads();
    const int cnt = ncnt;
    for (int i = w; i < cnt; i += NW) {
        const int b = nlist[i];
        float a, c, d;
        block256_tree3(V, G(V->soi_x), G(V->soi_y), G(V->mterm), b, a, c, d);
        if (lane == 0) { lv[b] = a; lv[nbpad + b] = c; lv[2 * nbpad + b] = d; }
    }
    __syncthreads();
    // totals, in every wave
    const float tx = tree(0), ty = tree(1);
    const float tm = tree(2);
    float scale = 1.0f, fx = tx, fy = ty;
    if (P.has_thrust_cap) {
        const float F_tot = sqrtf(tx * tx + ty * ty);
        scale = (F_tot > P.max_ship_ice_force) ? (P.max_ship_ice_force / vmax(F_tot, 1e-6f)) : 1.0f;
    }
    // with the cap binding, the totals of the scaled forces (x * 1 = x: at
    // scale 1 they are the totals above, bit for bit)
    if (scale != 1.0f) {
        __syncthreads();   // every wave has read the unscaled partials
        for (int i = w; i < cnt; i += NW) {
            const int b = nlist[i];
            const float a = block256_tree(V, G(V->soi_x), b, scale, true);
            const float c = block256_tree(V, G(V->soi_y), b, scale, true);
            if (lane == 0) { lv[b] = a; lv[nbpad + b] = c; }
        }
        __syncthreads();
        fx = tree(0);
        fy = tree(1);
    }
    if (w == 0) {
        if (threadIdx.x == 0) {
            G(V->scal)[0] = scale;
            ShipPre pre;
            if (preship) ship_pre_get(lpre, pre);
            ship_frame(V, D, k, last, fx, fy, tm, preship, pre);
            store_pose(V, D->s.d_ship);
            *near_count(V) = 0;
            *cand_count(V, set) = 0;   // consumed by the previous k_post, refilled by this frame's
        }
    } else {
        floe_eligibility(V, D, scale, set, 1);
        // frames that built the lists, counted beside the ship step
        if (threadIdx.x == 64 && V->nl_on && nl_rebuild(V, k)) atomicAdd(nl_rebuilds(V), 1u);
    }
}

// ── k_force ──────────────────────────────────────────────────────────────────

// per-element values staged for the block's epilogue (force_epilogue): the
// element's own state and statics, its hull and drag inputs, the group's
// contact and bond sums, and the flags (act_raw | act_bond << 1 | crack << 2)
enum : int { S_PX, S_PY, S_VX, S_VY, S_RO, S_OM, S_R, S_H, S_AREA, S_MASS, S_INERTIA, S_SUB0, S_PEN, S_FMR, S_SX,
             S_SY, S_NXE, S_NYE, S_VWX, S_VWY, S_FIX, S_FIY, S_MI, S_FBX, S_FBY, S_MB, S_FLAGS, S_N };
// Per group, one row of CFS floats holds the partners' x, y, vx, vy and
// radius (MAXC2 each) during the contact pass; once the group's lane 0 has
// summed the contact terms the row is dead and its first S_N floats stage
// the element's epilogue.  Sharing the row keeps the 8-lane block at
// 19.5 KB of LDS, 8 blocks per CU.  CFS = 84 = 20 mod 32: the 8 groups of a
// wave start on distinct banks.
constexpr int CFS = 5 * MAXC2 + 4;
static_assert(S_N <= CFS, "epilogue staging must fit a partner row");

template <int GR>
struct ForceLds {
    static constexpr int E = BLOCK / GR;
    int ck[E][MAXC2];                    // partner's rotated index rot_key(j, e, n)
    float cf[E][CFS];
    int ord[E][MAXC2];
    uint32_t grec[E][GR];                // scan-list cell's record offset
    int goff[E][GR];                     // its first slot in the group's slot order
    float bf[E][3][GR];
    __device__ float* cx(int g) { return cf[g]; }
    __device__ float* cy(int g) { return cf[g] + MAXC2; }
    __device__ float* cvx(int g) { return cf[g] + 2 * MAXC2; }
    __device__ float* cvy(int g) { return cf[g] + 3 * MAXC2; }
    __device__ float* cr(int g) { return cf[g] + 4 * MAXC2; }
    __device__ float* stg(int g) { return cf[g]; }
};

// extra (tail) workgroups of a frame's first k_force
constexpr int TAIL_GROUPS = 1;

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// group-wide inclusive prefix sum over GR lanes (16 = one DPP row; shifts
// into a row bring 0, and for 8-lane groups lanes below the shift take 0)
template <int GR>
__device__ __forceinline__ int group_incl_scan(int v) {
    const int l = threadIdx.x & (GR - 1);
    int t = __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);   // row_shr:1
    v += (GR == 16 || l >= 1) ? t : 0;
    t = __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);       // row_shr:2
    v += (GR == 16 || l >= 2) ? t : 0;
    t = __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);       // row_shr:4
    v += (GR == 16 || l >= 4) ? t : 0;
    if (GR == 16) v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);   // row_shr:8
    return v;
}

// value of the group's last lane: ds_swizzle bit-mask mode within each
// 32-lane half, source lane (l & and_mask) | or_mask
template <int GR>
__device__ __forceinline__ int group_last(int v) {
    constexpr int pat = (0x1F & ~(GR - 1)) | ((GR - 1) << 5);
    return __builtin_amdgcn_ds_swizzle(v, pat);
}

// The element's scan list: the fine cells [x0..x1] x [y0..y1] first, then
// the coarse cells [u0..u1] x [v0..v1]; cell q of the list is bucket
// record list_rec(q).
struct ScanList {
    int x0, y0, W, nfine;
    int u0, v0, W2, ntot;
    float iW, iW2;   // approximate 1 / W, 1 / W2 (0 for an empty part)
};

// q = row * W + col for 0 <= q < 2^22, W >= 1, iW within a few ulp of 1 / W:
// the float estimate of the row is off by at most one, fixed by one step
__device__ __forceinline__ void split_rc(int q, int W, float iW, int& row, int& col) {
    row = (int)((float)q * iW);
    col = q - row * W;
    if (col < 0) { row -= 1; col += W; }
    else if (col >= W) { row += 1; col -= W; }
}

__device__ __forceinline__ int list_rec(const DevView* __restrict__ V, int set, const ScanList& S, int q) {
    const bool fine = q < S.nfine;
    int row, col;
    split_rc(fine ? q : q - S.nfine, fine ? S.W : S.W2, fine ? S.iW : S.iW2, row, col);
    return fine ? fine_rec(V, set, (S.y0 + row) * V->p.grid_nx + (S.x0 + col))
                : coarse_rec(V, set, (S.v0 + row) * V->nx2 + (S.u0 + col));
}

// The occupied slots of the scan list, GR cells at a time: the lanes read
// the cell counts, a group prefix sum lays the occupied slots out, and each
// lane takes one slot per round.  visit(have, j, kin, r) is called by every
// lane of the group, the same number of times.
template <int GR, class F>
__device__ __forceinline__ void scan_list(const DevView* __restrict__ V, int set, const ScanList& S,
                                          ForceLds<GR>& L, int g, F&& visit) {
    const int l = threadIdx.x & (GR - 1);
    for (int c0 = 0; c0 < S.ntot; c0 += GR) {
        const int q = c0 + l;
        uint32_t off = 0;
        int cc = 0;
        if (q < S.ntot) {
            const int rix = list_rec(V, set, S, q);
            off = rec_off(V, rix);
            cc = min(*(GAS const int*)(G(V->cells) + off), rec_cap(V, rix));
        }
        const int incl = group_incl_scan<GR>(cc);
        const int total = group_last<GR>(incl);
        L.grec[g][l] = off;
        L.goff[g][l] = incl - cc;   // lanes past the list: total
        wave_sync_lds();
        for (int base = 0; base < total; base += GR) {
            // slot f lies in the last cell qq whose first slot is <= f (cells
            // without slots share their successor's offset; offsets past the
            // list are total > f), at position f - goff[qq] < its count
            const int f = base + l;
            const bool have = f < total;
            Slot sl = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, -1};
            if (have) {
                int qq = 0;
                #pragma unroll
                for (int t = 1; t < GR; ++t) qq += (L.goff[g][t] <= f) ? 1 : 0;
                sl = slot_load(G(V->cells) + L.grec[g][qq], f - L.goff[g][qq]);
            }
            visit(have, sl.e, make_float4(sl.x, sl.y, sl.vx, sl.vy), sl.r);
        }
        wave_sync_lds();
    }
}

// Ice-ice contact force on element e (group-uniform), summed in the
// reference's scatter order; result valid in the group's lane 0.
// (px, py, pvx, pvy, pr) is e's own state; partners come from bucket set
// `set` (and, when buckets overflowed, from pos/vel/radius).
template <int GR>
__device__ __forceinline__ void contact_force_group(const DevView* __restrict__ V, int set, GAS const float* __restrict__ pos,
                                    GAS const float* __restrict__ vel, int e, float px, float py, float pvx,
                                    float pvy, float pr, int novf, ForceLds<GR>& L, int g,
                                    float& fx, float& fy, float& fm) {
    const IceDemParams& P = V->p;
    const int n = V->n;
    const int l = threadIdx.x & (GR - 1);
    const int gshift = (threadIdx.x & 63) & ~(GR - 1);
    int count = 0;
    auto push = [&](bool ov, int j, float x, float y, float vx, float vy, float r) {
        const uint64_t bal = __ballot(ov);
        const uint32_t m = (uint32_t)(bal >> gshift) & ((1u << GR) - 1u);
        const int before = __builtin_popcount(m & ((1u << l) - 1u));
        if (ov && count + before < MAXC2) {
            const int s = count + before;
            L.ck[g][s] = rot_key(j, e, n);
            L.cx(g)[s] = x; L.cy(g)[s] = y; L.cvx(g)[s] = vx; L.cvy(g)[s] = vy; L.cr(g)[s] = r;
        }
        count += __builtin_popcount(m);
    };
    // the pair's operands in (lo, hi) order by selects, so that one inlined
    // copy of the pair arithmetic serves both orders
    auto visit = [&](bool have, int j, float4 k4, float r) {
        const bool lo = e < j;
        const bool ov = have && j != e &&
                        overlap_v(P, lo ? px : k4.x, lo ? py : k4.y, lo ? pr : r, lo ? k4.x : px, lo ? k4.y : py,
                                  lo ? r : pr);
        push(ov, j, k4.x, k4.y, k4.z, k4.w, r);
    };
    int x0, x1, y0, y1, u0 = 0, u1 = -1, v0 = 0, v1 = -1;
    scan_range(P, pr, px, py, x0, x1, y0, y1);
    if (V->nbig > 0) scan_range2(V, px, py, u0, u1, v0, v1);
    ScanList S;
    S.x0 = x0; S.y0 = y0; S.W = x1 - x0 + 1; S.nfine = S.W * (y1 - y0 + 1);
    S.u0 = u0; S.v0 = v0; S.W2 = u1 - u0 + 1; S.ntot = S.nfine + S.W2 * (v1 - v0 + 1);
    S.iW = __builtin_amdgcn_rcpf((float)S.W);
    S.iW2 = S.W2 > 0 ? __builtin_amdgcn_rcpf((float)S.W2) : 0.0f;
    scan_list<GR>(V, set, S, L, g, visit);
    for (int base = 0; base < novf; base += GR) {
        const int f = base + l;
        bool ov = false;
        int j = -1;
        float x = 0.0f, y = 0.0f, vx = 0.0f, vy = 0.0f, r = 0.0f;
        if (f < novf) {
            j = G(V->ovf_list)[(size_t)set * V->ovf_cap + f];
            x = pos[j]; y = pos[n + j]; vx = vel[j]; vy = vel[n + j]; r = G(V->radius)[j];
            const bool lo = e < j;
            ov = j != e && overlap_v(P, lo ? px : x, lo ? py : y, lo ? pr : r, lo ? x : px, lo ? y : py, lo ? r : pr);
        }
        push(ov, j, x, y, vx, vy, r);
    }
    fx = 0.0f; fy = 0.0f; fm = 0.0f;
    if (count <= MAXC2) {
        // every partner was seen once: rank by rotated partner index
        for (int i = l; i < count; i += GR) {
            const int ki = L.ck[g][i];
            int rank = 0;
            for (int q = 0; q < count; ++q) rank += (L.ck[g][q] < ki) ? 1 : 0;
            L.ord[g][rank] = i;
        }
        wave_sync_lds();
        // contributions in rank order, written over the partner arrays once
        // every lane has read its inputs; lane 0 adds them in order
        constexpr int PER = (MAXC2 + GR - 1) / GR;
        float ca[PER], cb[PER], cc[PER];
        #pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int i = l + u * GR;
            ca[u] = 0.0f; cb[u] = 0.0f; cc[u] = 0.0f;
            if (i < count) {
                const int s = L.ord[g][i];
                const float x = L.cx(g)[s], y = L.cy(g)[s], vx = L.cvx(g)[s], vy = L.cvy(g)[s], r = L.cr(g)[s];
                const bool lo = L.ck[g][s] < n - e;   // e < j
                contrib_v(P, lo, lo ? px : x, lo ? py : y, lo ? pvx : vx, lo ? pvy : vy, lo ? pr : r,
                          lo ? x : px, lo ? y : py, lo ? vx : pvx, lo ? vy : pvy, lo ? r : pr, ca[u], cb[u], cc[u]);
            }
        }
        wave_sync_lds();
        #pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int i = l + u * GR;
            if (i < count) { L.cx(g)[i] = ca[u]; L.cy(g)[i] = cb[u]; L.cvx(g)[i] = cc[u]; }
        }
        wave_sync_lds();
        if (l == 0)
            for (int i = 0; i < count; ++i) { fx = fx + L.cx(g)[i]; fy = fy + L.cy(g)[i]; fm = fm + L.cvx(g)[i]; }
        wave_sync_lds();
        return;
    }
    // more than MAXC2 contacts: exact selection of the next partner by lane 0
    if (l == 0) {
        int last = 0;
        for (;;) {
            int best = 0x7fffffff, bj = -1;
            auto consider = [&](int j) {
                if (j == e) return;
                const int kj = rot_key(j, e, n);
                if (kj <= last || kj >= best) return;
                if (!pair_overlap(P, pos, G(V->radius), n, min(e, j), max(e, j))) return;
                best = kj; bj = j;
            };
            for (int gy = y0; gy <= y1; ++gy)
                for (int gx = x0; gx <= x1; ++gx) {
                    GAS const uint8_t* rec = rec_ptr(V, fine_rec(V, set, gy * P.grid_nx + gx));
                    const int cc = min(*(GAS const int*)rec, P.grid_bucket);
                    for (int s = 0; s < cc; ++s) consider(slot_load(rec, s).e);
                }
            for (int gy = v0; gy <= v1; ++gy)
                for (int gx = u0; gx <= u1; ++gx) {
                    GAS const uint8_t* rec = rec_ptr(V, coarse_rec(V, set, gy * V->nx2 + gx));
                    const int cc = min(*(GAS const int*)rec, V->cap2);
                    for (int s = 0; s < cc; ++s) consider(slot_load(rec, s).e);
                }
            for (int s = 0; s < novf; ++s) consider(G(V->ovf_list)[(size_t)set * V->ovf_cap + s]);
            if (bj < 0) break;
            last = best;
            float a, b, c;
            pair_contrib(P, pos, vel, G(V->radius), n, min(e, bj), max(e, bj), e < bj, a, b, c);
            fx = fx + a; fy = fy + b; fm = fm + c;
        }
    }
}

// ── partner lists ────────────────────────────────────────────────────────────

// In-order running sum across a lane group: lane l adds its term c after the
// terms of lanes 0 .. l - 1, starting from the group-uniform acc (a DPP chain
// of single-lane row shifts).  Returns the new group-uniform total.  The
// order is the sequential one of the reference's scatter-add; a +0 term
// leaves a sum unchanged (sums that start at +0 never reach -0).
template <int GR>
__device__ __forceinline__ float group_chain(float acc, float c) {
    const int l = threadIdx.x & (GR - 1);
    float s = (l == 0) ? acc + c : c;
    #pragma unroll
    for (int t = 1; t < GR; ++t) {
        const float prev = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(s), 0x111, 0xf, 0xf, false));
        s = (l == t) ? prev + c : s;
    }
    return __int_as_float(group_last<GR>(__float_as_int(s)));
}

// element j's slot (state at the start of the frame) for the list pass
__device__ __forceinline__ Slot es_load(const DevView* __restrict__ V, int j) {
    GAS const float* q = G(V->es) + 6 * (size_t)j;
    const f4u a = *(GAS const f4u*)q;
    Slot o;
    o.x = a.x; o.y = a.y; o.vx = a.z; o.vy = a.w; o.r = q[4]; o.e = j;
    return o;
}

// list candidate: centre distance below r_i + r_j + 2 delta + skin (a
// superset of the contacts of every frame until the next build)
__device__ __forceinline__ bool within_skin(const IceDemParams& P, float skin, float xa, float ya, float ra,
                                            float xb_, float yb_, float rb) {
    const float dx = xb_ - xa, dy = yb_ - ya;
    float md = ra + rb;
    if (P.contact_dilation2 != 0.0f) md = md + P.contact_dilation2;
    return sqrtf(dx * dx + dy * dy) < md + skin;
}

// Ice-ice contact force on element e from its partner list (group-uniform;
// result valid in the group's lane 0; with sums false only the list is
// built, on a build frame, and the result is 0).  On a build frame the grids (reach
// extended by the skin) give the candidates, ranked by rotated index into the
// list; an element with more than NLW candidates is marked and takes
// contact_force_group, this frame and every later one of the call.  The sum
// runs over the list in the reference's scatter order with +0 for listed
// partners out of contact, so it equals the sum over the contacts.
template <int GR>
__device__ __forceinline__ void list_force_group(const DevView* __restrict__ V, int set, GAS const float* __restrict__ pos,
                                                 GAS const float* __restrict__ vel, int e, float px, float py,
                                                 float pvx, float pvy, float pr, int novf, bool rebuild,
                                                 int cnt_in, bool sums, ForceLds<GR>& L, int g, float& fx, float& fy,
                                                 float& fm) {
    const IceDemParams& P = V->p;
    const int n = V->n;
    const int l = threadIdx.x & (GR - 1);
    const int gshift = (threadIdx.x & 63) & ~(GR - 1);
    int* lk = (int*)L.cf[g];          // candidate keys during a build (the row is free until the epilogue)
    int* lj = lk + NLW;               // the built list
    GAS int32_t* nlrow = G(V->nl) + (size_t)e * NLW;
    int cnt;
    if (rebuild) {
        int count = 0;
        const float skin = V->nl_skin;
        auto push = [&](bool c, int j) {
            const uint64_t bal = __ballot(c);
            const uint32_t m = (uint32_t)(bal >> gshift) & ((1u << GR) - 1u);
            const int before = __builtin_popcount(m & ((1u << l) - 1u));
            if (c && count + before < NLW) lk[count + before] = rot_key(j, e, n);
            count += __builtin_popcount(m);
        };
        auto visit = [&](bool have, int j, float4 k4, float r) {
            push(have && j != e && within_skin(P, skin, px, py, pr, k4.x, k4.y, r), j);
        };
        int x0, x1, y0, y1, u0 = 0, u1 = -1, v0 = 0, v1 = -1;
        cell_box(P, px, py, big_reach(P, pr) + skin, x0, x1, y0, y1);
        if (V->nbig > 0) scan_range2(V, px, py, u0, u1, v0, v1);
        ScanList S;
        S.x0 = x0; S.y0 = y0; S.W = x1 - x0 + 1; S.nfine = S.W * (y1 - y0 + 1);
        S.u0 = u0; S.v0 = v0; S.W2 = u1 - u0 + 1; S.ntot = S.nfine + S.W2 * (v1 - v0 + 1);
        S.iW = __builtin_amdgcn_rcpf((float)S.W);
        S.iW2 = S.W2 > 0 ? __builtin_amdgcn_rcpf((float)S.W2) : 0.0f;
        scan_list<GR>(V, set, S, L, g, visit);
        for (int base = 0; base < novf; base += GR) {
            const int f = base + l;
            bool c = false;
            int j = -1;
            if (f < novf) {
                j = G(V->ovf_list)[(size_t)set * V->ovf_cap + f];
                c = j != e && within_skin(P, skin, px, py, pr, pos[j], pos[n + j], G(V->radius)[j]);
            }
            push(c, j);
        }
        wave_sync_lds();
        if (count > NLW) {
            if (l == 0) { G(V->nl_cnt)[e] = NLW + 1; atomicOr(nl_ovf(V), 1u); }
            wave_sync_lds();
            if (sums) contact_force_group<GR>(V, set, pos, vel, e, px, py, pvx, pvy, pr, novf, L, g, fx, fy, fm);
            else { fx = 0.0f; fy = 0.0f; fm = 0.0f; }
            return;
        }
        // rank by rotated index: the reference's order (pairs (e, j > e) by j,
        // then (i < e, e) by i)
        for (int i = l; i < count; i += GR) {
            const int ki = lk[i];
            int rank = 0;
            for (int q = 0; q < count; ++q) rank += (lk[q] < ki) ? 1 : 0;
            const int j = (ki < n - e) ? e + ki : e + ki - n;
            lj[rank] = j;
            nlrow[rank] = j;
        }
        if (l == 0) G(V->nl_cnt)[e] = count;
        wave_sync_lds();
        if (!sums) { fx = 0.0f; fy = 0.0f; fm = 0.0f; return; }   // list built, forces not needed
        cnt = count;
    } else {
        cnt = cnt_in;
        if (cnt > NLW) {
            contact_force_group<GR>(V, set, pos, vel, e, px, py, pvx, pvy, pr, novf, L, g, fx, fy, fm);
            return;
        }
    }
    // 8 lanes: two rounds of partners per pass, both rounds' list entries,
    // then both rounds' slots, in flight together (16 lanes: one round; the
    // second set of registers cost occupancy there)
    constexpr int NR = GR < 16 ? 2 : 1;
    float ax = 0.0f, ay = 0.0f, am = 0.0f;
    auto term = [&](bool have, const Slot& sl, float& cx, float& cy, float& cm) {
        cx = 0.0f; cy = 0.0f; cm = 0.0f;
        if (!have) return;
        const bool lo = e < sl.e;
        const float xl = lo ? px : sl.x, yl = lo ? py : sl.y, rl = lo ? pr : sl.r;
        const float xh = lo ? sl.x : px, yh = lo ? sl.y : py, rh = lo ? sl.r : pr;
        if (overlap_v(P, xl, yl, rl, xh, yh, rh))
            contrib_v(P, lo, xl, yl, lo ? pvx : sl.vx, lo ? pvy : sl.vy, rl, xh, yh, lo ? sl.vx : pvx,
                      lo ? sl.vy : pvy, rh, cx, cy, cm);
    };
    for (int base = 0; base < cnt; base += NR * GR) {
        Slot sl[NR];
        bool h[NR];
        #pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int q = base + r * GR + l;
            h[r] = q < cnt;
            sl[r].e = h[r] ? (rebuild ? lj[q] : nlrow[q]) : 0;
        }
        #pragma unroll
        for (int r = 0; r < NR; ++r)
            if (h[r]) sl[r] = es_load(V, sl[r].e);
        #pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (base + r * GR < cnt) {
                float cx, cy, cm;
                term(h[r], sl[r], cx, cy, cm);
                ax = group_chain<GR>(ax, cx);
                ay = group_chain<GR>(ay, cy);
                am = group_chain<GR>(am, cm);
            }
        }
    }
    if (rebuild) wave_sync_lds();     // lj is read above; the row is the epilogue's staging next
    fx = ax; fy = ay; fm = am;
}

// Partner and bond state of one bond-table entry (the dynamic reads of the
// bond pass).
struct BondIn {
    float ox, oy, dmg;
    bool intact, act;
};

__device__ __forceinline__ BondIn bond_in(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                          GAS const float* __restrict__ pos, int4 t4, int mode, int set) {
    BondIn o;
    o.ox = pos[t4.y]; o.oy = pos[V->n + t4.y];
    o.intact = G(D->s.d_bond_intact)[t4.x] != 0;
    o.dmg = G(D->s.d_bond_damage)[t4.x];
    o.act = G(D->s.d_active)[t4.y] != 0;
    if (mode != 1) o.act = o.act || (G(floe_coll_of(V, set))[t4.z] != 0u);
    return o;
}

// Bond forces on element e (jax_simulation.py:63-108): lanes compute the terms
// of the element's bond-table entries, lane 0 adds them in table order (the
// order of the reference's two scatter-add passes).  act_e is e's activity
// as the bond pass sees it; (t0, p0, in0) is the lane's entry l with its
// state, read ahead by the caller.  Result valid in lane 0.
template <int GR>
__device__ __forceinline__ void bond_force_group(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                                 GAS const float* __restrict__ pos, int e, float px, float py,
                                                 bool act_e, int deg, int4 t0, float2 p0, BondIn in0, int mode,
                                                 int set, ForceLds<GR>& L, int g, float& fbx, float& fby, float& mb,
                                                 bool& crack) {
    const int l = threadIdx.x & (GR - 1);
    const int gshift = (threadIdx.x & 63) & ~(GR - 1);
    fbx = 0.0f; fby = 0.0f; mb = 0.0f;
    crack = false;
    for (int base = 0; base < deg; base += GR) {
        const int q = base + l;
        bool broken = false;
        if (q < deg) {
            int4 t4 = t0;
            float2 pr = p0;
            BondIn bi = in0;
            if (base > 0) {
                t4 = ldg(G(V->btab) + (size_t)e * V->W + q);
                pr = ldg(G(V->bprm) + (size_t)e * V->W + q);
                bi = bond_in(V, D, pos, t4, mode, set);
            }
            const bool role_j = t4.w != 0;
            broken = !bi.intact;
            const float xi = role_j ? bi.ox : px, yi = role_j ? bi.oy : py;
            const float xj = role_j ? px : bi.ox, yj = role_j ? py : bi.oy;
            const bool ai = role_j ? bi.act : act_e, aj = role_j ? act_e : bi.act;
            float dx = xj - xi;
            float dy = yj - yi;
            float dist = sqrtf(dx * dx + dy * dy);
            float dsafe = vmax(dist, 1e-10f);
            float dirx = dx / dsafe, diry = dy / dsafe;
            float stretch = dist - pr.x;
            float Fm = pr.y * stretch * (1.0f - bi.dmg);
            float Fx = dirx * Fm, Fy = diry * Fm;
            if (!bi.intact) { Fx = 0.0f; Fy = 0.0f; }
            float ok = ((ai && aj) || (!ai && !aj)) ? 1.0f : 0.0f;
            Fx = Fx * ok; Fy = Fy * ok;
            float cross = 0.5f * (dx * Fy - dy * Fx);
            L.bf[g][0][l] = role_j ? -Fx : Fx;
            L.bf[g][1][l] = role_j ? -Fy : Fy;
            L.bf[g][2][l] = role_j ? -cross : cross;
        }
        const uint64_t bal = __ballot(broken);
        crack = crack || (((uint32_t)(bal >> gshift) & ((1u << GR) - 1u)) != 0u);
        wave_sync_lds();
        if (l == 0) {
            const int m = min(GR, deg - base);
            for (int t = 0; t < m; ++t) {
                fbx = fbx + L.bf[g][0][t]; fby = fby + L.bf[g][1][t]; mb = mb + L.bf[g][2][t];
            }
        }
        wave_sync_lds();
    }
}

// One bond-table entry with its bond force formed by the previous frame's
// k_post, and the partner's activity as the bond pass sees it.
struct BondPre {
    float fx, fy, cr;
    bool intact, act;
};

__device__ __forceinline__ BondPre bond_pre(const DevView* __restrict__ V, const CallDesc* __restrict__ D, int4 t4,
                                            int mode, int set) {
    const f4u f = *(GAS const f4u*)(G(V->bfv) + t4.x);
    BondPre o;
    o.fx = f.x; o.fy = f.y; o.cr = f.z; o.intact = f.w != 0.0f;
    o.act = G(D->s.d_active)[t4.y] != 0;
    if (mode != 1) o.act = o.act || (G(floe_coll_of(V, set))[t4.z] != 0u);
    return o;
}

// Bond forces on element e from the forces k_post formed once per bond
// (frames after a call's first, one ice sub-step): each table entry keeps
// its bond's force when both ends are active or both frozen, negated for the
// j role, and the sums run in table order (jax_simulation.py:93-106).  A
// dropped entry would contribute a signed zero in the reference's
// arithmetic; +0 terms leave the ordered sums unchanged.  Result valid in
// every lane of the group.
template <int GR>
__device__ __forceinline__ void bond_pre_group(const DevView* __restrict__ V, const CallDesc* __restrict__ D, int e,
                                               bool act_e, int deg, int4 t0, BondPre b0, int mode, int set,
                                               float& fbx, float& fby, float& mb, bool& crack) {
    const int l = threadIdx.x & (GR - 1);
    const int gshift = (threadIdx.x & 63) & ~(GR - 1);
    float ax = 0.0f, ay = 0.0f, am = 0.0f;
    crack = false;
    for (int base = 0; base < deg; base += GR) {
        const int q = base + l;
        float cx = 0.0f, cy = 0.0f, cm = 0.0f;
        bool broken = false;
        if (q < deg) {
            int4 t4 = t0;
            BondPre b = b0;
            if (base > 0) {
                t4 = ldg(G(V->btab) + (size_t)e * V->W + q);
                b = bond_pre(V, D, t4, mode, set);
            }
            const bool role_j = t4.w != 0;
            broken = !b.intact;
            const bool ai = role_j ? b.act : act_e, aj = role_j ? act_e : b.act;
            if ((ai && aj) || (!ai && !aj)) {
                cx = role_j ? -b.fx : b.fx; cy = role_j ? -b.fy : b.fy; cm = role_j ? -b.cr : b.cr;
            }
        }
        const uint64_t bal = __ballot(broken);
        crack = crack || (((uint32_t)(bal >> gshift) & ((1u << GR) - 1u)) != 0u);
        ax = group_chain<GR>(ax, cx);
        ay = group_chain<GR>(ay, cy);
        am = group_chain<GR>(am, cm);
    }
    fbx = ax; fby = ay; mb = am;
}

// Crack flag of a frozen element: whether any of its bond-table entries is
// broken (the only output of the bond pass that a frozen element's frame
// keeps).  intact0 is the lane's entry l as the caller read it; pre: the
// flags come from the records k_post formed, else from the bond state.
template <int GR>
__device__ __forceinline__ bool crack_group(const DevView* __restrict__ V, const CallDesc* __restrict__ D, int e,
                                            int deg, bool intact0, bool pre) {
    const int l = threadIdx.x & (GR - 1);
    const int gshift = (threadIdx.x & 63) & ~(GR - 1);
    bool crack = false;
    for (int base = 0; base < deg; base += GR) {
        const int q = base + l;
        bool broken = false;
        if (q < deg) {
            bool it = intact0;
            if (base > 0) {
                const int4 t4 = ldg(G(V->btab) + (size_t)e * V->W + q);
                it = pre ? (((GAS const f4u*)G(V->bfv))[t4.x].w != 0.0f) : (G(D->s.d_bond_intact)[t4.x] != 0);
            }
            broken = !it;
        }
        const uint64_t bal = __ballot(broken);
        crack = crack || (((uint32_t)(bal >> gshift) & ((1u << GR) - 1u)) != 0u);
    }
    return crack;
}

// The epilogue of an element that stays frozen in frame k > 0 of a call with
// one ice sub-step (mode 0).  Its next state, force magnitude (0: a hull
// contact would have activated it), activity and list-build position are
// what frame 0 wrote, so only its submergence (the no-contact branch of
// force_epilogue's update, which needs no thickness) and its crack flag are
// formed.  Returns the squared list displacement, 0.
__device__ __forceinline__ bool frozen_fast(const IceDemParams& P, int k, int sub, int mode) {
    return mode == 0 && sub == 0 && k > 0 && P.n_ice_steps == 1;
}
__device__ __forceinline__ float frozen_epilogue(const DevView* __restrict__ V, const CallDesc* __restrict__ D, int e,
                                                 bool crack) {
    const IceDemParams& P = V->p;
    const float sub0 = G(D->s.d_sub)[e];
    G(D->s.d_sub)[e] = vmax(0.0f, sub0 - P.inv_rise_time * sub0 * P.dt);
    G(V->crack)[e] = crack ? 1 : 0;
    return 0.0f;
}

// The epilogue of an element frozen through a call's first frame (mode 0,
// one ice sub-step): what force_epilogue's masked integration leaves, formed
// without its other inputs.  The next state is the state (px, py, vx, vy,
// ro, om); the force magnitude fmr is the near test's (not positive: a
// positive one needs a hull contact, which would have activated the
// element's floe); the submergence takes the no-contact branch (likewise);
// the activity stays 0, so it is not written.  With lists it returns the
// squared displacement from the build position (0 on a build frame, which
// records the position).
__device__ __forceinline__ bool frozen_first(const IceDemParams& P, int k, int sub, int mode) {
    return mode == 0 && sub == 0 && k == 0 && P.n_ice_steps == 1;
}
__device__ __forceinline__ float frozen_first_epilogue(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                                       int e, float px, float py, float vx, float vy, float ro,
                                                       float om, float sub0, float fmr, bool lists, bool rebuild) {
    const IceDemParams& P = V->p;
    const int n = V->n;
    float bx = px, by = py;
    if (lists && !rebuild) { bx = G(V->xb)[e]; by = G(V->xb)[n + e]; }
    G(V->nxt_pos)[e] = px; G(V->nxt_pos)[n + e] = py;
    G(V->nxt_vel)[e] = vx; G(V->nxt_vel)[n + e] = vy;
    G(V->nxt_rot)[e] = ro; G(V->nxt_omega)[e] = om;
    G(V->fm)[e] = fmr;
    G(D->s.d_sub)[e] = vmax(0.0f, sub0 - P.inv_rise_time * sub0 * P.dt);
    if (!lists) return 0.0f;
    if (rebuild) { G(V->xb)[e] = px; G(V->xb)[n + e] = py; }
    const float ddx = px - bx, ddy = py - by;
    return ddx * ddx + ddy * ddy;
}

// The epilogue of element e by one thread (jax_simulation.py:140-268): ship
// forces and submergence (first sub-step), drag, rotational drag, totals and
// the integration into the next-state buffers, from the values its group
// staged in sv.
// With partner lists it returns the squared displacement of the element's
// new position from its position at the last list build (written here on a
// build frame), else 0.
__device__ __forceinline__ float force_epilogue(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                                int sub, int src_is_state, int mode,
                                                const float* __restrict__ fship_in, float* __restrict__ fout,
                                                float* __restrict__ mout, const float* sv, int e,
                                                bool lists = false, bool rebuild = false) {
    const IceDemParams& P = V->p;
    const int n = V->n;
    GAS float* dpos = G(src_is_state ? V->nxt_pos : D->s.d_pos);
    GAS float* dvel = G(src_is_state ? V->nxt_vel : D->s.d_vel);
    GAS float* drot = G(src_is_state ? V->nxt_rot : D->s.d_rot);
    GAS float* domg = G(src_is_state ? V->nxt_omega : D->s.d_omega);
    const float px = sv[S_PX], py = sv[S_PY], vx = sv[S_VX], vy = sv[S_VY], ro = sv[S_RO], om = sv[S_OM];
    const float r = sv[S_R], h = sv[S_H], area = sv[S_AREA], mass = sv[S_MASS], inertia = sv[S_INERTIA];
    const float sub0 = sv[S_SUB0], pen = sv[S_PEN], fmr = sv[S_FMR], sx_ = sv[S_SX], sy_ = sv[S_SY];
    const float nxe = sv[S_NXE], nye = sv[S_NYE], vwx = sv[S_VWX], vwy = sv[S_VWY];
    const float fix = sv[S_FIX], fiy = sv[S_FIY], mi = sv[S_MI], fbx = sv[S_FBX], fby = sv[S_FBY], mb = sv[S_MB];
    const int fl = __float_as_int(sv[S_FLAGS]);
    const bool act_raw = (fl & 1) != 0, act_bond = (fl & 2) != 0, crack = (fl & 4) != 0;
    float shx, shy, subv;
    bool act;
    if (mode == 1) {
        act = act_raw;
        shx = fship_in[e]; shy = fship_in[n + e];
        subv = sub0;
    } else if (sub == 0) {
        act = act_bond;
        G(D->s.d_active)[e] = act ? 1 : 0;
        G(V->fm)[e] = fmr;   // k_post applies the frame's thrust scale
        const bool hc = pen > 0.0f;
        const float pen_c = vmin(pen, P.submergence_pen_cap);
        float ns = hc ? vmin(sub0 + P.submergence_rate * pen_c * P.dt, h)
                      : vmax(0.0f, sub0 - P.inv_rise_time * sub0 * P.dt);
        G(D->s.d_sub)[e] = ns;
        subv = ns;
        float z = vmin(ns, h);
        float Fb = P.buoyancy_coef * area * z;
        shx = sx_ + (hc ? Fb * nxe : 0.0f);
        shy = sy_ + (hc ? Fb * nye : 0.0f);
        if (P.n_ice_steps > 1) { G(V->shipf)[e] = shx; G(V->shipf)[n + e] = shy; }
    } else {
        act = act_raw;
        shx = G(V->shipf)[e]; shy = G(V->shipf)[n + e];
        subv = sub0;
    }
    if (sub == 0 && mode == 0) G(V->crack)[e] = crack ? 1 : 0;
    // drag (:167-192)
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
    float A_sub = area * vmin(subv / vmax(h, 1e-10f), 1.0f);
    float r3 = r * (r * r);
    float Md = P.rot_drag_coef * r3 * om * fabsf(om) * A_sub;
    const float Fx_t = ((fbx + fix) - fdx) + shx;
    const float Fy_t = ((fby + fiy) - fdy) + shy;
    const float M_t = (mb + mi) + Md;
    if (mode == 1) {
        fout[e] = Fx_t; fout[n + e] = Fy_t; mout[e] = M_t;
        return 0.0f;
    }
    // integration (:215-268)
    const float dt = P.dt_ice;
    const float vc = P.velocity_clip, pc = P.position_clip;
    float ax = Fx_t / mass, ay = Fy_t / mass;
    float nvx = vclip(vx + ax * dt, -vc, vc);
    float nvy = vclip(vy + ay * dt, -vc, vc);
    float npx = px + nvx * dt;
    float npy = py + nvy * dt;
    float alpha = M_t / inertia;
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
    float d2 = 0.0f;
    if (lists) {
        float bx, by;
        if (rebuild) { bx = px; by = py; G(V->xb)[e] = px; G(V->xb)[n + e] = py; }
        else { bx = G(V->xb)[e]; by = G(V->xb)[n + e]; }
        const float ddx = npx - bx, ddy = npy - by;
        d2 = ddx * ddx + ddy * ddy;
        if (act) {   // stamps of a moving element (see stamped_near)
            const int cn = coarse_idx(V, npx, npy), cb = coarse_idx(V, bx, by);
            if (rebuild || !act_raw || cn != cb) {
                const uint32_t st = *G(nl_rebuilds(V)) + 1u;
                G(V->dstamp)[cn] = st;
                if (rebuild) G(V->dstamp)[cb] = st;
            }
        }
    }
    return d2;
}


// elements [chunk*E, chunk*E+E), E = 256 / GR, GR lanes each: the lane
// groups form the contact and bond sums and stage them with the element's
// values; the epilogues run on the group's lane 0 (GR = 16) or on threads
// [0, E) after a block barrier (GR = 8; every thread of the block calls;
// `again`: the block calls once more, so the staging area must be free on
// return).  mode 0: a frame sub-step (integrates); mode 1: forces out
// (fout, mout).
// The epilogue's inputs that neither the contact nor the bond pass reads:
// rotation and spin, statics, the sub-step accumulator, the hull terms of a
// near element (first sub-step of a frame only) and the water velocity.
struct EpiIn {
    float ro, om, h, area, mass, inertia, sub0, pen, fmr, sx, sy, nxe, nye, vwx, vwy;
};

__device__ __forceinline__ EpiIn epi_in(const DevView* __restrict__ V, const CallDesc* __restrict__ D, int k, int sub,
                                        int src_is_state, int mode, int e) {
    EpiIn o;
    o.ro = G(src_is_state ? D->s.d_rot : V->nxt_rot)[e];
    o.om = G(src_is_state ? D->s.d_omega : V->nxt_omega)[e];
    o.h = G(V->thick)[e]; o.area = G(V->area)[e]; o.mass = G(V->mass)[e]; o.inertia = G(V->inertia)[e];
    o.sub0 = G(D->s.d_sub)[e];
    o.pen = 0.0f; o.fmr = 0.0f; o.sx = 0.0f; o.sy = 0.0f; o.nxe = 0.0f; o.nye = 0.0f;
    if (mode == 0 && sub == 0 && G(V->nearf)[e] != 0) {
        o.pen = G(V->pen)[e]; o.fmr = G(V->fm_raw)[e]; o.sx = G(V->soi_x)[e]; o.sy = G(V->soi_y)[e];
        o.nxe = G(V->nx)[e]; o.nye = G(V->ny)[e];
    }
    o.vwx = 0.0f; o.vwy = 0.0f;
    GAS const float* vwp = G(D->in.d_v_water);
    if (vwp) {
        const int64_t base = (int64_t)k * D->in.vw_stride_k + (int64_t)e * D->in.vw_stride_e;
        o.vwx = vwp[base];
        o.vwy = vwp[base + D->in.vw_stride_c];
    }
    return o;
}

__device__ __forceinline__ void epi_put(float* sv, const EpiIn& o) {
    sv[S_RO] = o.ro; sv[S_OM] = o.om; sv[S_H] = o.h; sv[S_AREA] = o.area; sv[S_MASS] = o.mass;
    sv[S_INERTIA] = o.inertia; sv[S_SUB0] = o.sub0; sv[S_PEN] = o.pen; sv[S_FMR] = o.fmr; sv[S_SX] = o.sx;
    sv[S_SY] = o.sy; sv[S_NXE] = o.nxe; sv[S_NYE] = o.nye; sv[S_VWX] = o.vwx; sv[S_VWY] = o.vwy;
}

template <int GR>
__device__ __forceinline__ void force_body(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                           int k, int sub, int src_is_state, int mode, int set,
                           const float* __restrict__ fship_in, float* __restrict__ fout, float* __restrict__ mout,
                           ForceLds<GR>& L, int chunk, bool again = false, bool frames = false) {
    const int n = V->n;
    const int g = threadIdx.x / GR, l = threadIdx.x & (GR - 1);
    const int e = chunk * epb<GR>() + g;
    // frames-mode kernels (a k_post between frames): partner lists, and bond
    // forces formed by the previous frame's k_post
    const bool lists = frames && V->nl_on != 0 && mode == 0;
    const bool pre = frames && V->bf_on != 0 && mode == 0 && k > 0;
    const bool rebuild = lists && nl_rebuild(V, k);
    float d2 = 0.0f;
    if (e < n) {                              // group-uniform
        GAS const float* pos = G(src_is_state ? D->s.d_pos : V->nxt_pos);
        GAS const float* vel = G(src_is_state ? D->s.d_vel : V->nxt_vel);
        GAS const uint8_t* active = G(D->s.d_active);
        // everything that depends only on e, read in one round trip
        const float px = pos[e], py = pos[n + e], vx = vel[e], vy = vel[n + e];
        const float r = G(V->radius)[e];
        const int fl = G(V->floe)[e];
        const bool act_raw = active[e] != 0;
        const int deg = V->W > 0 ? G(V->bdeg)[e] : 0;
        const int novf = G(V->ovf_count)[set];
        const int lcnt = (lists && !rebuild) ? G(V->nl_cnt)[e] : 0;   // read with the element
        int4 bt0 = make_int4(0, 0, 0, 0);
        float2 bp0 = make_float2(0.0f, 0.0f);
        if (l < V->W) { bt0 = ldg(G(V->btab) + (size_t)e * V->W + l); bp0 = ldg(G(V->bprm) + (size_t)e * V->W + l); }
        // 16 lanes: the epilogue-only inputs too, in the same round trip
        EpiIn ei;
        if (GR >= 16) ei = epi_in(V, D, k, sub, src_is_state, mode, e);
        const bool fcoll = (mode != 1) && (G(floe_coll_of(V, set))[fl] != 0u);
        const bool act_bond = act_raw || fcoll;   // activity seen by the bond pass (mode 0)
        BondIn bin0 = {0.0f, 0.0f, 0.0f, false, false};
        BondPre bpre0 = {0.0f, 0.0f, 0.0f, false, false};
        // read ahead of the contact pass for an element known to be live
        // (active, or every element in mode 1); with 8 lanes (large fields)
        // one that its floe's contact activates reads its bond state after
        // the contact pass and a frozen one none; with 16 lanes (small
        // fields, latency-bound) every element reads ahead
        const bool ahead = mode != 0 || act_raw || GR >= 16;
        if (ahead && l < deg) {
            if (pre) bpre0 = bond_pre(V, D, bt0, mode, set);
            else bin0 = bond_in(V, D, pos, bt0, mode, set);
        }
        // A frame sub-step leaves an element that is frozen in it where it
        // is whatever its forces (the integration is masked by elem_active,
        // jax_simulation.py:256-261), so a frozen element forms no contact or
        // bond sums and no crack flag (the fracture pass reads an element's
        // crack flag only for bonds whose ends are both active): on a build
        // frame only its list.
        const bool live = mode != 0 || (sub == 0 ? act_bond : act_raw);
        float fix = 0.0f, fiy = 0.0f, mi = 0.0f;
        if (lists) {
            if (live || rebuild)
                list_force_group<GR>(V, set, pos, vel, e, px, py, vx, vy, r, novf, rebuild, lcnt, live, L, g, fix,
                                     fiy, mi);
        } else if (live) {
            contact_force_group<GR>(V, set, pos, vel, e, px, py, vx, vy, r, novf, L, g, fix, fiy, mi);
        }
        float fbx = 0.0f, fby = 0.0f, mb = 0.0f;
        bool crack = false;
        if (live && !ahead && l < deg) {
            if (pre) bpre0 = bond_pre(V, D, bt0, mode, set);
            else bin0 = bond_in(V, D, pos, bt0, mode, set);
        }
        if (!live) {
        } else if (pre) {
            bond_pre_group<GR>(V, D, e, act_bond, deg, bt0, bpre0, mode, set, fbx, fby, mb, crack);
        } else {
            bond_force_group<GR>(V, D, pos, e, px, py, mode == 1 ? act_raw : act_bond, deg, bt0, bp0, bin0, mode, set, L,
                                 g, fbx, fby, mb, crack);
        }
        if (l == 0) {
            float* sv = L.stg(g);
            sv[S_PX] = px; sv[S_PY] = py; sv[S_VX] = vx; sv[S_VY] = vy; sv[S_R] = r;
            if (GR >= 16) epi_put(sv, ei);
            sv[S_FIX] = fix; sv[S_FIY] = fiy; sv[S_MI] = mi; sv[S_FBX] = fbx; sv[S_FBY] = fby; sv[S_MB] = mb;
            sv[S_FLAGS] = __int_as_float((act_raw ? 1 : 0) | (act_bond ? 2 : 0) | (crack ? 4 : 0));
            // 16 lanes (small fields, latency-bound): the group's lane 0
            // finishes its own element, no barrier
            if (GR >= 16) {
                if (!live && frozen_fast(V->p, k, sub, mode))
                    d2 = frozen_epilogue(V, D, e, crack);
                else if (!live && frozen_first(V->p, k, sub, mode))
                    d2 = frozen_first_epilogue(V, D, e, px, py, vx, vy, ei.ro, ei.om, ei.sub0, ei.fmr, lists,
                                               rebuild);
                else
                    d2 = force_epilogue(V, D, sub, src_is_state, mode, fship_in, fout, mout, sv, e, lists, rebuild);
            }
        }
    }
    if (GR < 16) {
        // 8 lanes (large fields, instruction-bound): one thread per element
        // of the block, so the epilogue's instructions serve 32 elements
        __syncthreads();
        const int et = chunk * epb<GR>() + (int)threadIdx.x;
        if ((int)threadIdx.x < epb<GR>() && et < n) {
            // the epilogue-only inputs, loaded here (coalesced, one thread per
            // element) rather than held in registers through the passes
            float* sv = L.stg(threadIdx.x);
            const int fl = __float_as_int(sv[S_FLAGS]);
            if ((fl & 2) == 0 && frozen_fast(V->p, k, sub, mode)) {
                d2 = frozen_epilogue(V, D, et, (fl & 4) != 0);
            } else if ((fl & 2) == 0 && frozen_first(V->p, k, sub, mode)) {
                const float fmr = G(V->nearf)[et] != 0 ? G(V->fm_raw)[et] : 0.0f;
                d2 = frozen_first_epilogue(V, D, et, sv[S_PX], sv[S_PY], sv[S_VX], sv[S_VY], G(D->s.d_rot)[et],
                                           G(D->s.d_omega)[et], G(D->s.d_sub)[et], fmr, lists, rebuild);
            } else {
                epi_put(sv, epi_in(V, D, k, sub, src_is_state, mode, et));
                d2 = force_epilogue(V, D, sub, src_is_state, mode, fship_in, fout, mout, sv, et, lists, rebuild);
            }
        }
    }
    if (lists) {
        // frame k + 1 rebuilds when a displacement passes the threshold: a
        // wave whose largest one does stores it into its slot (any value past
        // the threshold will do, so racing stores need no read-modify-write);
        // the other waves store nothing
        const float m = wave_max(d2);
        if ((threadIdx.x & 63) == 0 && m > V->nl_thr2)
            __hip_atomic_store(nl_slot(V, k + 1) + (blockIdx.x & (NLSUB - 1)), __float_as_uint(m), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
    if (GR < 16 && again) __syncthreads();
}

// The force pass of a frame after a call's first, one ice sub-step
// (frozen_fast), for elements [chunk*E, chunk*E+E).  An element frozen
// through the frame (elem_active false and no hull contact on its floe this
// frame) keeps its state whatever its forces (jax_simulation.py:256-261), so
// it reads only its activity, floe and submergence: no contact or bond sums,
// no next state (k_post keeps a frozen element's state), no force magnitude
// (0 since the call's first frame: a contact would have activated it), and no
// crack flag (k_post reads it only for bonds with both ends active).  A live
// element reads what force_body reads, one round trip later; on a list-build
// frame a frozen element also builds its list.
template <int GR>
__device__ __forceinline__ void force_body_lazy(const DevView* __restrict__ V, const CallDesc* __restrict__ D, int k,
                                                int set, ForceLds<GR>& L, int chunk) {
    const int n = V->n;
    const int g = threadIdx.x / GR, l = threadIdx.x & (GR - 1);
    const int e = chunk * epb<GR>() + g;
    const bool lists = V->nl_on != 0;
    const bool pre = V->bf_on != 0;
    const bool rebuild = lists && nl_rebuild(V, k);
    float d2 = 0.0f;
    if (e < n) {                              // group-uniform
        GAS const float* pos = G(D->s.d_pos);
        GAS const float* vel = G(D->s.d_vel);
        const bool act_raw = G(D->s.d_active)[e] != 0;
        const int fl = G(V->floe)[e];
        const float sub0 = G(D->s.d_sub)[e];
        const bool need = rebuild && G(V->rneed)[e] != 0;   // read with the element
        float px = 0.0f, py = 0.0f, vx = 0.0f, vy = 0.0f, r = 0.0f;
        int deg = 0, novf = 0, lcnt = 0;
        int4 bt0 = make_int4(0, 0, 0, 0);
        float2 bp0 = make_float2(0.0f, 0.0f);
        EpiIn ei;
        auto load_elem = [&]() {
            px = pos[e]; py = pos[n + e]; vx = vel[e]; vy = vel[n + e];
            // 16 lanes: the epilogue-only inputs too, in the same round trip
            if (GR >= 16) ei = epi_in(V, D, k, 0, 1, 0, e);
            r = G(V->radius)[e];
            deg = V->W > 0 ? G(V->bdeg)[e] : 0;
            novf = G(V->ovf_count)[set];
            lcnt = (lists && !rebuild) ? G(V->nl_cnt)[e] : 0;
            if (l < V->W) {
                bt0 = ldg(G(V->btab) + (size_t)e * V->W + l);
                if (!pre) bp0 = ldg(G(V->bprm) + (size_t)e * V->W + l);
            }
        };
        // an active element's state is read beside nothing else; a frozen
        // one first learns whether its floe was hit this frame
        if (act_raw) load_elem();
        const bool fcoll = !act_raw && G(floe_coll_of(V, set))[fl] != 0u;
        const bool live = act_raw || fcoll;
        // on a build frame a frozen element rebuilds its list only when the
        // frame before marked it (partial builds, stamped_near)
        const bool rb = rebuild && (live || need);
        if (!act_raw && (live || rb)) load_elem();
        BondIn bin0 = {0.0f, 0.0f, 0.0f, false, false};
        BondPre bpre0 = {0.0f, 0.0f, 0.0f, false, false};
        if (live && l < deg) {
            if (pre) bpre0 = bond_pre(V, D, bt0, 0, set);
            else bin0 = bond_in(V, D, pos, bt0, 0, set);
        }
        float fix = 0.0f, fiy = 0.0f, mi = 0.0f;
        if (lists) {
            if (live || rb)
                list_force_group<GR>(V, set, pos, vel, e, px, py, vx, vy, r, novf, rb, lcnt, live, L, g, fix,
                                     fiy, mi);
        } else if (live) {
            contact_force_group<GR>(V, set, pos, vel, e, px, py, vx, vy, r, novf, L, g, fix, fiy, mi);
        }
        float fbx = 0.0f, fby = 0.0f, mb = 0.0f;
        bool crack = false;
        if (live) {
            if (pre)
                bond_pre_group<GR>(V, D, e, true, deg, bt0, bpre0, 0, set, fbx, fby, mb, crack);
            else
                bond_force_group<GR>(V, D, pos, e, px, py, true, deg, bt0, bp0, bin0, 0, set, L, g, fbx, fby, mb,
                                     crack);
        }
        if (l == 0) {
            float* sv = L.stg(g);
            if (live) {
                sv[S_PX] = px; sv[S_PY] = py; sv[S_VX] = vx; sv[S_VY] = vy; sv[S_R] = r;
                sv[S_FIX] = fix; sv[S_FIY] = fiy; sv[S_MI] = mi; sv[S_FBX] = fbx; sv[S_FBY] = fby; sv[S_MB] = mb;
                sv[S_FLAGS] = __int_as_float((act_raw ? 1 : 0) | 2 | (crack ? 4 : 0));
                if (GR >= 16) {
                    epi_put(sv, ei);
                    d2 = force_epilogue(V, D, 0, 1, 0, nullptr, nullptr, nullptr, sv, e, lists, rebuild);
                }
            } else {
                // the no-contact branch of the submergence update; +0 stays
                // +0 bit for bit, so only a positive depth is stored
                if (__float_as_uint(sub0) != 0u)
                    G(D->s.d_sub)[e] = vmax(0.0f, sub0 - V->p.inv_rise_time * sub0 * V->p.dt);
                // its list was built where it stands
                if (rb) { G(V->xb)[e] = px; G(V->xb)[n + e] = py; }
                sv[S_FLAGS] = __int_as_float(8);   // done
            }
        }
    }
    if (GR < 16) {
        __syncthreads();
        const int et = chunk * epb<GR>() + (int)threadIdx.x;
        if ((int)threadIdx.x < epb<GR>() && et < n) {
            float* sv = L.stg(threadIdx.x);
            if ((__float_as_int(sv[S_FLAGS]) & 8) == 0) {
                epi_put(sv, epi_in(V, D, k, 0, 1, 0, et));
                d2 = force_epilogue(V, D, 0, 1, 0, nullptr, nullptr, nullptr, sv, et, lists, rebuild);
            }
        }
    }
    if (lists) {
        const float m = wave_max(d2);
        if ((threadIdx.x & 63) == 0 && m > V->nl_thr2)
            __hip_atomic_store(nl_slot(V, k + 1) + (blockIdx.x & (NLSUB - 1)), __float_as_uint(m), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
}

// With nextra > 0 (a frame's first sub-step) blocks [0, nextra) run
// force_tail (dispatched first, so the ship step overlaps the element
// work); blocks [nextra, nextra + fblk) integrate their elements.
// Workgroup ids are dealt to the 8 XCDs round-robin; chunk_of maps them so
// that each XCD takes one contiguous run of element chunks (its L2 then
// serves the bucket and partner reads of spatially close elements when the
// element order is spatial).  A bijection on [0, total).
__device__ __forceinline__ int chunk_of(int w, int total) {
    constexpr int NX = 8;
    const int q = total / NX, r = total % NX;
    const int x = w % NX, slot = w / NX;
    return (x < r) ? x * (q + 1) + slot : r * (q + 1) + (x - r) * q + slot;
}

// workgroup lb of lg = nextra + element chunks of one context's force pass;
// xcd: lb is the hardware workgroup id (chunks spread XCD by XCD)
template <int GR>
__device__ __forceinline__ void force_kernel_body(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                                  int k, int sub, int src_is_state, int mode, int set, int last,
                                                  int nextra, int lb, int lg, bool xcd,
                                                  const float* __restrict__ fship_in, float* __restrict__ fout,
                                                  float* __restrict__ mout, ForceLds<GR>& L,
                                                  bool preship = false) {
    if (lb < nextra) {
        force_tail(V, D, k, last, set, preship, (float*)&L, (int)sizeof(ForceLds<GR>));
        return;
    }
    const int nf = lg - nextra;
    const int chunk = xcd ? chunk_of(lb - nextra, nf) : lb - nextra;
    if (frozen_fast(V->p, k, sub, mode))
        force_body_lazy<GR>(V, D, k, set, L, chunk);
    else
        force_body<GR>(V, D, k, sub, src_is_state, mode, set, fship_in, fout, mout, L, chunk, false, true);
}

template <int GR>
__global__ __launch_bounds__(BLOCK) void k_force(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                                 int k, int sub, int src_is_state, int mode, int set, int last,
                                                 int nextra,
                                                 const float* __restrict__ fship_in,
                                                 float* __restrict__ fout, float* __restrict__ mout, int preship) {
    __shared__ ForceLds<GR> L;
    force_kernel_body<GR>(V, D, k, sub, src_is_state, mode, set, last, nextra, (int)blockIdx.x, (int)gridDim.x, true,
                      fship_in, fout, mout, L, preship != 0);
}

// ── k_post: fracture (jax_simulation.py:534-656), commit, histories, next cull ─

// contact force magnitude of element i after the frame's thrust scale
// (jax_simulation.py:456-470): fm holds it unscaled, scal[0] the scale
__device__ __forceinline__ float fm_scaled(const DevView* __restrict__ V, int i) {
    const float f = G(V->fm)[i];
    return V->p.has_thrust_cap ? f * G(V->scal)[0] : f;
}

// Bond t's fracture update (the DPDEM gates and collision breaks of
// jax_simulation.py:534-656 around update_bond_damage_state,
// jax_bond_fracture.py:134).  Reads only what k_post does not
// write (final positions and velocities, activity, eligibility, force
// magnitudes), so it runs beside the element work in any order.
__device__ __forceinline__ void post_bond(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                          int k, int final_in_nxt, int t, bool prebond = false) {
    const IceDemParams& P = V->p;
    const int n = V->n, nb = V->nb;
    GAS const float* pos = G(final_in_nxt ? V->nxt_pos : D->s.d_pos);
    GAS const float* vel = G(final_in_nxt ? V->nxt_vel : D->s.d_vel);
    if (t < nb) {
        const int i = G(V->bi)[t], j = G(V->bj)[t];
        GAS const uint8_t* act = G(D->s.d_active);
        if (act[i] == 0 && act[j] == 0) {
            // Both ends frozen after the frame's activation: neither moved,
            // neither has a hull contact (a positive force magnitude needs
            // one, and a contact activates the element's floe), so every
            // gate below is closed and the bond's state comes out as it went
            // in (jax_simulation.py:563-640 select the old values; the shear
            // sum adds +0, which turns only a -0 into +0).  Activity never
            // reverts, so both ends have been frozen since the call's first
            // frame, whose k_post formed the force record from the same
            // positions and state: only that frame writes it, and only that
            // frame can find the -0 the shear sum turns into +0 (a later
            // frame of the call finds what the first one left).
            const bool intact = G(D->s.d_bond_intact)[t] != 0;
            if (D->h.d_bond_intact) G(D->h.d_bond_intact)[(size_t)k * nb + t] = intact ? 1 : 0;
            if (k != 0) return;
            GAS float* dsp = G(D->s.d_bond_ds) + t;
            if (__float_as_uint(*dsp) == 0x80000000u) *dsp = 0.0f;
            if (!prebond) return;
        }
        float dx = pos[j] - pos[i];
        float dy = pos[n + j] - pos[n + i];
        float dist = sqrtf(dx * dx + dy * dy);
        float ds_ = vmax(dist, 1e-10f);
        float dirx = dx / ds_, diry = dy / ds_;
        GAS const uint8_t* elig = G(D->s.d_floe_elig);
        bool both = act[i] && act[j];
        const int2 bf = ldg(G(V->bfloe) + t);
        bool floe_ok = elig[bf.x] && elig[bf.y & 0x3fffffff];
        bool breakable = (bf.y >> 30) == 0;
        const float fmi = fm_scaled(V, i), fmj = fm_scaled(V, j);
        bool near = (fmi > 0.0f) || (fmj > 0.0f);
        bool at_crack = G(V->crack)[i] || G(V->crack)[j];
        bool dpdem_ok = both && floe_ok && (near || at_crack) && breakable;
        float vrx = vel[j] - vel[i];
        float vry = vel[n + j] - vel[n + i];
        float vn = vrx * dirx + vry * diry;
        float vtx = vrx - vn * dirx, vty = vry - vn * diry;
        float ds_incr = sqrtf(vtx * vtx + vty * vty) * P.dt_ice;
        bool gate = near && both && floe_ok && breakable;
        const float ds_old = G(D->s.d_bond_ds)[t];
        float new_ds = ds_old + (gate ? ds_incr : 0.0f);
        const float rest = G(V->rest)[t];
        float L0 = vmax(rest, 1e-10f);
        float delta_n = dist - rest;
        float dn_d = near ? delta_n : 0.0f;
        float sn_el = P.k_n * (dn_d / L0);
        float st_el = P.k_s * (new_ds / L0);
        const float dmg = G(D->s.d_bond_damage)[t];
        float sigma_n = sn_el * (1.0f - dmg);
        float sigma_t = st_el * (1.0f - dmg);
        const float dmx = G(D->s.d_bond_dmmax)[t], dm0 = G(D->s.d_bond_dm0)[t];
        const bool intact = G(D->s.d_bond_intact)[t] != 0;
        // the damage update only where its results are used (every one of
        // them is selected by dpdem_ok below)
        float nd = dmg, ndmx = dmx, ndm0 = dm0;
        bool nint = intact;
        if (dpdem_ok)
            bond_damage_update(delta_n, new_ds, sigma_n, sigma_t, G(V->sbar)[t], P.bond_cohesion, P.bond_friction,
                               P.G_c_I, P.G_c_II, P.eta_bk, dmg, dmx, dm0, intact, nd, ndmx, ndm0, nint);
        float out_d = dpdem_ok ? nd : dmg;
        float out_mx = dpdem_ok ? ndmx : dmx;
        float out_m0 = dpdem_ok ? ndm0 : dm0;
        bool int_d = (dpdem_ok && breakable) ? nint : intact;
        float cf = vmax(fmi, fmj);
        float eff = G(V->bstr)[t] * P.eff_strength_collision;
        bool brk = intact && near && (cf > eff) && breakable;
        bool fin = int_d && !brk;
        G(D->s.d_bond_damage)[t] = out_d;
        G(D->s.d_bond_dmmax)[t] = out_mx;
        G(D->s.d_bond_dm0)[t] = out_m0;
        G(D->s.d_bond_ds)[t] = new_ds;
        G(D->s.d_bond_intact)[t] = fin ? 1 : 0;
        if (D->h.d_bond_intact) G(D->h.d_bond_intact)[(size_t)k * nb + t] = fin ? 1 : 0;
        if (prebond) {
            // the next frame's bond force (jax_simulation.py:76-106) from the
            // frame's final positions, damage and intact flag, with the float
            // operations of bond_force_group; the activity factor stays with
            // the force pass
            const float stretch = dist - rest;
            const float Fm = G(V->bstf)[t] * stretch * (1.0f - out_d);
            float Fx = dirx * Fm, Fy = diry * Fm;
            if (!fin) { Fx = 0.0f; Fy = 0.0f; }
            const float cross = 0.5f * (dx * Fy - dy * Fx);
            f4u o;
            o.x = Fx; o.y = Fy; o.z = cross; o.w = fin ? 1.0f : 0.0f;
            *(GAS f4u*)(G(V->bfv) + t) = o;
        }
    }
}

// Thread t of chunk t / 256 (every thread of the 256-thread block calls):
// element t, cell t of the set just used, floe t.  With `cull`, the
// element's committed state also goes through the next frame's near test
// and into bucket set set ^ 1.
// With partner lists the buckets are filled only for a frame that builds
// lists (or when an element overflowed its list this call) and cleared only
// after a frame that used them.
__device__ __forceinline__ void post_elem(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                          int k, int final_in_nxt, int last, int t, int chunk, int set, int cull,
                                          bool lists = false, NearLds* __restrict__ nl = nullptr) {
    const int n = V->n;
    GAS const float* pos = G(final_in_nxt ? V->nxt_pos : D->s.d_pos);
    GAS const float* vel = G(final_in_nxt ? V->nxt_vel : D->s.d_vel);
    float px = 0.0f, py = 0.0f, pvx = 0.0f, pvy = 0.0f, ro = 0.0f, r = 0.0f, bnd = 0.0f;
    // the list control words, read beside the element's state
    bool used = true, ins = true, built = false, nb1 = false;
    if (lists) {
        const bool ovf = nl_overflowed(V);
        // frame 0's buckets were filled by the call's cull whether or not the
        // frame built (and its call-start slot is cleared below, unread)
        built = k == 0 || nl_rebuild(V, k);
        used = ovf || built;
        nb1 = nl_rebuild(V, k + 1);
        ins = ovf || nb1;
    }
    // a frozen element's next state is a copy of its state (the integration
    // keeps it), and its partner-list slot is the one the call's first cull
    // wrote: neither is written again
    bool frozen = false, listed = false;
    if (t < n) {
        px = pos[t]; py = pos[n + t];
        pvx = vel[t]; pvy = vel[n + t];
        if (cull) { r = G(V->radius)[t]; bnd = G(V->bound)[t]; }   // statics for the cull, read with the state
        // with nl: whether the next frame's near candidates list the element
        if (cull && nl) listed = G(V->cflag)[(size_t)(set ^ 1) * n + t] != 0;
        frozen = G(D->s.d_active)[t] == 0;
        // the next frame builds the lists (after a call's first frame): this
        // element's list is rebuilt when it is active or a live element may
        // have come within reach of it
        if (lists && nb1) G(V->rneed)[t] = (!frozen || stamped_near(V, px, py, *G(nl_rebuilds(V)))) ? 1 : 0;
        if (final_in_nxt && !frozen) {
            G(D->s.d_pos)[t] = px; G(D->s.d_pos)[n + t] = py;
            G(D->s.d_vel)[t] = pvx; G(D->s.d_vel)[n + t] = pvy;
            G(D->s.d_rot)[t] = G(V->nxt_rot)[t]; G(D->s.d_omega)[t] = G(V->nxt_omega)[t];
        }
        if (D->h.d_positions) {
            G(D->h.d_positions)[(size_t)k * 2 * n + t] = px;
            G(D->h.d_positions)[(size_t)k * 2 * n + n + t] = py;
        }
        if (D->h.d_rotations || nl) ro = final_in_nxt ? G(V->nxt_rot)[t] : G(D->s.d_rot)[t];
        if (D->h.d_rotations) G(D->h.d_rotations)[(size_t)k * n + t] = ro;
        if (D->h.d_submergence) G(D->h.d_submergence)[(size_t)k * n + t] = G(D->s.d_sub)[t];
        if (last) {
            if (D->dg.d_pen_max) G(D->dg.d_pen_max)[t] = G(V->nearf)[t] ? G(V->pen)[t] : 0.0f;
            if (D->dg.d_force_mag) G(D->dg.d_force_mag)[t] = fm_scaled(V, t);
        }
    }
    if (lists) {
        if (t == 0) {
            #pragma unroll
            for (int i = 0; i < NLSUB; ++i) {
                __hip_atomic_store(nl_slot(V, k + 2) + i, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                // the call-start slot, read by frame 0's force pass, is
                // clear for the next call's cull
                if (k == 0) __hip_atomic_store(nl_start(V) + i, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            // the overflow flag is final once the force pass has run
            if (last) __hip_atomic_store(nl_ovf_prev(V), *nl_ovf(V), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if (used) clear_cells(V, set, t);
    if (t < V->nf) { G(floe_coll_of(V, set))[t] = 0u; G(floe_maxfm_of(V, set))[t] = 0u; }
    // after a call's last frame no candidate list is pending
    if (last && t == 0) { *cand_count(V, 0) = 0; *cand_count(V, 1) = 0; }
    if (cull && chunk < V->nblk) {
        bool near;
        (void)cull_chunk(V, chunk, t < n, t, px, py, pvx, pvy, r, bnd, G(D->s.d_ship), set ^ 1, near,
                         nl == nullptr, ins, !frozen, nl ? set : -1);
        if (nl) {
            // the candidates' near tests run in the kernel's near workgroups;
            // a near element they miss takes its test here
            const bool miss = near && !listed;
            const int cm = __syncthreads_count(miss ? 1 : 0);
            near_inline(V, D, nl, cm, miss, t, px, py, ro, set ^ 1, false);
        }
    }
}

// Near workgroup lb of lg of a single field's k_post: one wave per listed
// candidate of the next frame's near test (parity par), with the state this
// frame committed; a candidate passing the exact test (the cull's, with the
// same inputs) takes its hull test and pre-pass.  Runs from the kernel's
// start, beside the cull.
__device__ __forceinline__ void near_cand_block(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                                int lb, int lg, int final_in_nxt, int par) {
    const int n = V->n;
    const int wave = lb * (BLOCK / 64) + (threadIdx.x >> 6);
    const int nwaves = lg * (BLOCK / 64);
    GAS const int32_t* cl = G(V->cand) + (size_t)(par & 1) * n;
    const int count = *G(cand_count(V, par));
    const int e0 = (wave < n) ? cl[wave] : 0;   // read beside the count
    if (wave >= count) return;
    GAS const float* ship = G(D->s.d_ship);
    const ShipPose SP = stored_pose(V, D);
    GAS const float* pos = G(final_in_nxt ? V->nxt_pos : D->s.d_pos);
    GAS const float* rotp = G(final_in_nxt ? V->nxt_rot : D->s.d_rot);
    for (int q = wave; q < count; q += nwaves) {
        const int e = (q == wave) ? e0 : cl[q];
        // the hull test's statics read with the state, before the exact test
        const float x = pos[e], y = pos[n + e], ro = rotp[e], bnd = G(V->bound)[e];
        const NearIn ni = near_in(V, e);
        if (near_box(V->p, x, y, bnd, ship, 1.0f)) near_elem_at(V, SP, e, x, y, ro, par, ni);
    }
}

// Both parts for thread t (the one-workgroup scene kernel).
__device__ __forceinline__ void post_body(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                          int k, int final_in_nxt, int last, int t, int chunk, int set, int cull) {
    post_bond(V, D, k, final_in_nxt, t);
    post_elem(V, D, k, final_in_nxt, last, t, chunk, set, cull);
}

// Post workgroup b of a context's post grid: blocks [0, nbe) take elements,
// cells and floes, blocks [nbe, nbe + ceil(nb / 256)) take bonds, so neither
// chain waits behind the other in one thread and a small field spreads over
// more CUs.  xcd: b is the hardware workgroup id, and each part's chunks are
// spread XCD by XCD (chunk_of), so bonds and elements with neighbouring
// indices share an L2.
__device__ __forceinline__ void post_block(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                           int k, int final_in_nxt, int last, int b, int nbe, int set, int cull,
                                           int nbb, bool xcd, NearLds* __restrict__ nl) {
    if (b >= nbe) {
        const int c = xcd ? chunk_of(b - nbe, nbb) : b - nbe;
        post_bond(V, D, k, final_in_nxt, c * BLOCK + (int)threadIdx.x, V->bf_on != 0);
    } else {
        const int c = xcd ? chunk_of(b, nbe) : b;
        post_elem(V, D, k, final_in_nxt, last, c * BLOCK + (int)threadIdx.x, c, set, cull, V->nl_on != 0, nl);
    }
}

// shipblk: the grid's last workgroup forms frame k + 1's ship_pre_block;
// nnw: the first nnw workgroups (a multiple of 8, so the others keep their
// XCD dealing) run the next frame's near tests over its candidates
__global__ __launch_bounds__(BLOCK) void k_post(const DevView* __restrict__ V, const CallDesc* __restrict__ D,
                                                int k, int final_in_nxt, int last, int set, int cull, int nbe,
                                                int shipblk, int nnw) {
    __shared__ NearLds nl;
    if (shipblk && (int)blockIdx.x == (int)gridDim.x - 1) {
        ship_pre_block(V, D, k + 1);
        return;
    }
    if ((int)blockIdx.x < nnw) {
        near_cand_block(V, D, blockIdx.x, nnw, final_in_nxt, set ^ 1);
        return;
    }
    post_block(V, D, k, final_in_nxt, last, (int)blockIdx.x - nnw, nbe, set, cull,
               (int)gridDim.x - shipblk - nnw - nbe, true, &nl);
}

// ── batched frames: the same kernels over many contexts in one launch ───────
// map[blockIdx.x] = (instance, workgroup of that instance, its workgroup
// count, instance's nextra for kb_force / element-block count nbe for
// kb_post); the instances' contexts and call descriptors
// are views[i], descs[i].  Every instance runs the same frame k.

// a call's first cull: frame 0's near tests inside the culling workgroups,
// frame 1's near candidates listed
__global__ __launch_bounds__(BLOCK) void kb_cull(const DevView* __restrict__ views, const CallDesc* __restrict__ descs,
                                                 const int4* __restrict__ map) {
    __shared__ NearLds nl;
    const int4 m = map[blockIdx.x];
    cull_kernel_body(views + m.x, descs + m.x, m.y, &nl);
}


// 8 lanes: held to 8 waves per SIMD (<= 64 VGPRs and <= 100 SGPRs, the
// excess SGPRs spilled to VGPR lanes; 19.5 KB of LDS is 8 blocks per CU).
// Measured on the 256-patch batch: 68.8 us per frame against 72.5 at 7
// waves.  The single-field k_force<8> measured slower with the same bound
// and is left to the compiler.
template <int GR>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(GR < 16 ? 8 : 1)))
void kb_force(const DevView* __restrict__ views, const CallDesc* __restrict__ descs,
              const int4* __restrict__ map, int k, int last, int set) {
    __shared__ ForceLds<GR> L;
    const int4 m = map[blockIdx.x];
    force_kernel_body<GR>(views + m.x, descs + m.x, k, 0, 1, 0, set, last, m.w, m.y, m.z, false, nullptr, nullptr,
                      nullptr, L);
}

// map entries with m.w < 0 (first in the grid, so they start with the
// kernel): near workgroup m.y of m.z of instance m.x, the next frame's near
// tests over the candidates the previous cull listed (none after a call's
// last frame); the others: post workgroup m.y of m.z, m.w element blocks,
// whose culls list the candidates of the frame after and test a near
// element the candidates missed
__global__ __launch_bounds__(BLOCK) void kb_post(const DevView* __restrict__ views, const CallDesc* __restrict__ descs,
                                                 const int4* __restrict__ map, int k, int last, int set) {
    __shared__ NearLds nl;
    const int4 m = map[blockIdx.x];
    if (m.w < 0) {
        if (!last) near_cand_block(views + m.x, descs + m.x, m.y, m.z, 1, set ^ 1);
        return;
    }
    post_block(views + m.x, descs + m.x, k, 1, last, m.y, m.w, set, last ? 0 : 1, m.z - m.w, false, &nl);
}


// ── batched rollouts: one workgroup simulates one scene for `frames` frames ──

// Totals, thrust cap (jax_simulation.py:456-479), floe eligibility
// (:511-524) and the ship step, by one workgroup; la, lb hold nbpad floats.
__device__ __forceinline__ void mid_body(const DevView* __restrict__ V, const CallDesc* __restrict__ D, float* la, float* lb,
                         int k, int last, int set) {
    const IceDemParams& P = V->p;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int nb = V->nblk, nbpad = V->nbpad;
    for (int b = w; b < nb; b += nw) {
        float a = 0.0f, c = 0.0f, d = 0.0f;
        if (G(V->nearcnt)[b] > 0) {
            a = block256_tree(V, G(V->soi_x), b, 1.0f, false);
            c = block256_tree(V, G(V->soi_y), b, 1.0f, false);
            d = block256_tree(V, G(V->mterm), b, 1.0f, false);
        }
        if (lane == 0) { G(V->part1)[b] = a; G(V->part1)[nbpad + b] = c; G(V->part1)[2 * nbpad + b] = d; }
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
    float fx = tx, fy = ty;
    if (P.has_thrust_cap) {
        for (int b = w; b < nb; b += nw) {
            float a = 0.0f, c = 0.0f;
            if (G(V->nearcnt)[b] > 0) {
                a = block256_tree(V, G(V->soi_x), b, scale, true);
                c = block256_tree(V, G(V->soi_y), b, scale, true);
            }
            if (lane == 0) { G(V->part2)[b] = a; G(V->part2)[nbpad + b] = c; }
        }
        __syncthreads();
        fx = lds_tree_sum(V->part2, nb, nbpad, la, lb);
        fy = lds_tree_sum(V->part2 + nbpad, nb, nbpad, la, lb);
    }
    if (threadIdx.x == 0) G(V->scal)[0] = scale;
    floe_eligibility(V, D, scale, set);
    __syncthreads();
    if (threadIdx.x == 0) {
        ship_frame(V, D, k, last, fx, fy, tm);
        *near_count(V) = 0;
    }
    __syncthreads();
}

__global__ __launch_bounds__(BLOCK) void k_scene_rollouts(const DevView* __restrict__ views,
                                                          const CallDesc* __restrict__ descs, int frames) {
    __shared__ ForceLds<GRP> L;
    __shared__ float la[256], lb[256];
    const DevView* V = views + blockIdx.x;
    const CallDesc* D = descs + blockIdx.x;
    const int n = V->n, nblk = V->nblk;
    const int fch = (n + epb<GRP>() - 1) / epb<GRP>();
    const int post = max(max(V->nb, n), max(V->G, V->nf));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int c = 0; c < nblk; ++c) {
        const int e = c * BLOCK + threadIdx.x;
        const bool valid = e < n;
        float x = 0.0f, y = 0.0f, vx = 0.0f, vy = 0.0f;
        float r = 0.0f, bnd = 0.0f;
        if (valid) {
            x = G(D->s.d_pos)[e]; y = G(D->s.d_pos)[n + e]; vx = G(D->s.d_vel)[e]; vy = G(D->s.d_vel)[n + e];
            r = G(V->radius)[e]; bnd = G(V->bound)[e];
        }
        bool near;
        cull_chunk(V, c, valid, e, x, y, vx, vy, r, bnd, G(D->s.d_ship), 0, near, true);
    }
    __syncthreads();
    for (int k = 0; k < frames; ++k) {
        const int last = (k == frames - 1) ? 1 : 0;
        const int set = k & 1;
        const int count = __hip_atomic_load(near_count(V), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (count > 0) {
            const float* ship = D->s.d_ship;
            const ShipPose SP = make_pose(ship[0], ship[1], ship[2], V->p.half_len, V->p.half_beam);
            for (int q = w; q < count; q += BLOCK / 64) near_elem(V, D, SP, G(V->near_list)[q], set);
        }
        (void)lane;
        __syncthreads();
        mid_body(V, D, la, lb, k, last, set);
        for (int c = 0; c < fch; ++c)
            force_body<GRP>(V, D, k, 0, 1, 0, set, nullptr, nullptr, nullptr, L, c, true);
        __syncthreads();
        for (int base = 0; base < post; base += BLOCK)
            post_body(V, D, k, 1, last, base + threadIdx.x, base / BLOCK, set, last ? 0 : 1);
        __syncthreads();
    }
}

}  // namespace icedem

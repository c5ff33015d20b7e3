// icedem_internal.h -- device-side views shared by the DPDEM kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/icedem.h"
#include "detmath.h"

namespace icedem {

// Pointers held in DevView / CallDesc are generic; reading them through the
// global address space lets the compiler emit global (not flat) accesses,
// whose completion is tracked apart from LDS and scalar loads.
#define GAS __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ GAS T* G(T* p) { return (GAS T*)p; }

// loads of the HIP vector structs through a global pointer
typedef float gas_f2 __attribute__((ext_vector_type(2)));
typedef int gas_i4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float2 ldg(GAS const float2* p) {
    const gas_f2 v = *(GAS const gas_f2*)p;
    return make_float2(v.x, v.y);
}
typedef int gas_i2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ int2 ldg(GAS const int2* p) {
    const gas_i2 v = *(GAS const gas_i2*)p;
    return make_int2(v.x, v.y);
}
__device__ __forceinline__ int4 ldg(GAS const int4* p) {
    const gas_i4 v = *(GAS const gas_i4*)p;
    return make_int4(v.x, v.y, v.z, v.w);
}

constexpr int BLOCK = 256;          // threads per block of the per-element kernels
constexpr int MAX_PARTIALS = 4096;  // block partials reduced in one workgroup (n <= 1M)
constexpr int NLW = 32;             // partner-list width (partners per element)
constexpr int NLSUB = 8;            // sub-slots of each displacement maximum (one per XCD)
constexpr int NLCTL = 5 * NLSUB;      // partner-list control words (layout at DevView::nlctl)

// Static data, per-frame scratch and parameters of one context (device memory).
struct DevView {
    IceDemParams p;
    int32_t n, nb, nf, maxv, S, G, nblk, nbpad, ovf_cap;
    // static
    const float2* polys;      // (n*maxv) body-frame vertices
    const uint8_t* mask;      // (n*maxv)
    const float* mass; const float* inertia; const float* radius; const float* area;
    const float* thick; const float* bound;
    const int32_t* floe;
    const int32_t* bi; const int32_t* bj;
    // (n_bonds) per bond: floe of i, floe of j | 1 << 30 when the bond is not
    // breakable (a thickness above max_fracture_thickness)
    const int2* bfloe;
    const float* rest; const float* bstr; const float* sbar;
    const float2* hull;       // (S)
    // bonds of each element, element-major with fixed width W, in the
    // reference's scatter order (bonds where the element is i, ascending,
    // then bonds where it is j, ascending)
    int32_t W;
    const int32_t* bdeg;      // (n) entries of each element
    const int4* btab;         // (n W) bond, other element, floe of other, role (0 = i, 1 = j)
    const float2* bprm;       // (n W) rest length, stiffness E * h_avg * factor
    // per-frame scratch
    // (fm: contact force magnitude before the frame's thrust scale)
    float* pen; float* nx; float* ny; float* fm_raw; float* soi_x; float* soi_y; float* mterm; float* fm;
    uint8_t* nearf;           // (n) element passed the hull cull this frame
    int32_t* near_list;       // (n) queue of near elements
    int32_t* nearcnt;         // (nbpad) near elements of each 256-element block, 0 past nblk
    int32_t* counters;        // [0] near elements, [1], [2] near candidates (frame parity) ([3] unused)
    int32_t* cand;            // (2, n) near candidates of the next near test of each frame parity
    uint8_t* cflag;           // (2, n) element listed among them
    float cand_margin;        // candidate test margin (1 m + one frame's largest displacements)
    // bond forces formed once per bond by k_post for the next frame (one ice
    // sub-step per frame): (Fx, Fy, torque, intact) without the activity factor
    int32_t bf_on;
    float* bstf;              // (nb) bond stiffness E * mean thickness * bond_stiffness_factor
    float4* bfv;              // (nb) the next frame's bond forces
    // partner lists (params list_skin > 0 and one ice sub-step per frame)
    int32_t nl_on;            // lists in use
    float nl_thr2;            // rebuild once an element may have moved sqrt(nl_thr2) since the build
    float nl_skin;            // list reach beyond r_i + r_j + 2 delta
    int32_t* nl;              // (n, NLW) partners in the order of the reference's scatter (rotated index)
    int32_t* nl_cnt;          // (n) listed partners; NLW + 1: more than NLW, the element scans the grids
    float* xb;                // (2, n) positions at the last build
    uint32_t* dstamp;         // (G2) per coarse cell: 1 + the build count when a live element last stood in it
    uint8_t* rneed;           // (n) the next build frame rebuilds the element's list (written by the frame before)
    float dreach;             // a list's reach: 2 r_max + contact dilation + skin (and a rounding margin)
    float icell2;             // 1 / cell2
    float* es;                // (n, 6) x, y, vx, vy, radius (and a pad) of every element for the frame
    uint32_t* nlctl;          // [0, 24) squared-displacement maxima (float bits) of frame k's positions in
                              // slots [8 (k % 3), 8 (k % 3) + 8); [24] a list overflowed this call;
                              // [25] frames rebuilt since creation; [26] [24] at the end of the last
                              // call; [32, 40) the call-start flags (+inf: frame 0 builds), set by the
                              // call's cull, cleared by its first k_post
    int32_t lanes;            // lanes per element in k_force (8 or 16)
    float* shipf;             // (2n) ship forces on ice, kept for sub-steps > 0
    uint8_t* crack;           // (n) element touches a broken bond
    // Two grids, each with two bucket sets (frame parity).  A bucket is one
    // record: a 16-byte header whose first int is the insertion count, then
    // up to cap slots of 24 bytes (x, y, vx, vy, radius, element index), the
    // record padded to a multiple of 128 bytes so a lightly filled bucket is
    // one cache line.  Record r: fine set s cell c is r = s G + c; coarse set
    // s cell c is r = 2 G + s G2 + c.  The fine grid (params grid_*) holds
    // the elements of radius <= grid_rsmall; the coarse grid (cell2 >=
    // 2 (2 r_max + delta), same origin) the larger ones.
    uint8_t* cells;
    int64_t rec_f, rec_c;     // record bytes, fine and coarse
    int64_t coarse_off;       // byte offset of the first coarse record
    int32_t nbig, nx2, ny2, G2, cap2;
    float cell2;
    int32_t* ovf_list;        // (2 ovf_cap) insertions that found their bucket full
    int32_t* ovf_count;       // (2)
    uint32_t* floe_coll; uint32_t* floe_maxfm;                      // (2, nf): frame parity
    float* part1;             // (3 * nbpad) block partials: sum soi_x, soi_y, moment
    float* part2;             // (2 * nbpad) block partials: sum of scaled soi
    float* scal;              // [0] thrust scale of the frame, [8..11] sin/cos of -heading and heading
    float* shpre;             // (32) the force-independent part of the frame's ship step
    float* nxt_pos; float* nxt_vel; float* nxt_rot; float* nxt_omega;
    int32_t* err;
};

// Caller buffers of one icedem_frames call (device memory, rewritten per call).
struct CallDesc {
    IceDemState s;
    IceDemInputs in;
    IceDemHistory h;
    IceDemDiag dg;
};

}  // namespace icedem

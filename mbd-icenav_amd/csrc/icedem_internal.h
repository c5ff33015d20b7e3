// icedem_internal.h -- device-side views shared by the DPDEM kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/icedem.h"
#include "detmath.h"

namespace icedem {

constexpr int BLOCK = 256;          // threads per block of the per-element kernels
constexpr int MAX_PARTIALS = 4096;  // block partials reduced in one workgroup (n <= 1M)

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
    const float* rest; const float* bstr; const float* sbar;
    const float2* hull;       // (S)
    const int32_t* csr_start; // (n+1) bonds of each element: i-role then j-role
    const int32_t* csr_ent;   // (2 nb) entry = bond << 1 | role (0 = i, 1 = j)
    // per-frame scratch
    float* pen; float* nx; float* ny; float* fm_raw; float* soi_x; float* soi_y; float* mterm; float* fm;
    uint8_t* nearf;           // (n) element passed the hull cull this frame
    int32_t* near_list;       // (n) queue of near elements
    int32_t* near_count;      // (1)
    uint8_t* nearblk;         // (nblk) block of 256 elements holds a near element
    float* shipf;             // (2n) ship forces on ice, kept for sub-steps > 0
    uint8_t* crack;           // (n) element touches a broken bond
    int32_t* cell_count;      // (G) elements inserted per cell this frame
    int32_t* cell_slots;      // (G * grid_bucket) bucket contents
    int32_t* ovf_list;        // (ovf_cap) insertions that found their bucket full
    int32_t* ovf_count;       // (1)
    int32_t* ticket;          // (1) workgroups of k_near that have finished
    uint32_t* floe_coll; uint32_t* floe_maxfm;                      // (nf)
    float* part1;             // (3 * nbpad) block partials: sum soi_x, soi_y, moment
    float* part2;             // (2 * nbpad) block partials: sum of scaled soi
    float* scal;              // [0] thrust scale, [1] M_ice_on_ship, [2..3] sum of scaled ship forces
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

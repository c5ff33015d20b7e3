// surrogate.hip -- kernels of the rollout-cost surrogate (gfx950).
//
// GEMMs run on the matrix cores: v_mfma_f32_16x16x32_bf16, bf16 operands,
// fp32 accumulation.  A workgroup of 4 wavefronts computes a 64 x 64 tile of
// C; each wave owns a 32 x 32 quarter as 2 x 2 MFMA tiles.  K advances 32 at
// a time through two LDS buffers (the next step's global loads in flight
// during the current step's MFMAs): A as (m, k) rows and B as (n, k) rows,
// so every MFMA operand fragment (8 consecutive k of one row or column) is
// one 16-byte LDS read.  Split-K writes fp32 partials that a second kernel adds in split
// order.  No floating-point atomics anywhere.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include "../../include/icedem.h"
#include "../../include/icesur.h"

void icedem_set_error(const std::string& msg);   // icedem.hip

namespace icesur {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x0040u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

constexpr int TM = 64, TN = 64, TK = 32, KP = TK + 8;   // LDS rows padded to 80 bytes

// 8 consecutive bf16 of a tile row or column from global memory: one 16-byte
// load where the run is in range and aligned, else element by element with
// zeros outside [0, lim) (lim counts from the run's first element)
__device__ __forceinline__ uint4 fetch8(const uint16_t* __restrict__ p, bool row_ok, int lim) {
    if (row_ok && lim >= 8 && ((uintptr_t)p & 15) == 0) return *reinterpret_cast<const uint4*>(p);
    uint16_t v[8];
    #pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (row_ok && j < lim) ? p[j] : (uint16_t)0;
    return make_uint4(v[0] | ((uint32_t)v[1] << 16), v[2] | ((uint32_t)v[3] << 16),
                      v[4] | ((uint32_t)v[5] << 16), v[6] | ((uint32_t)v[7] << 16));
}
__device__ __forceinline__ uint16_t lane16(const uint4& q, int j) {
    const uint32_t w = j < 2 ? q.x : j < 4 ? q.y : j < 6 ? q.z : q.w;
    return (uint16_t)((j & 1) ? (w >> 16) : (w & 0xffffu));
}

// this thread's share of the op(A) tile [m0, m0 + 64) x [k0, k0 + 32) (stored
// as As[m][k]): 8 consecutive k of one row (A is M x K) or 8 consecutive m
// of one k (A is K x M)
template <bool TA>
__device__ __forceinline__ uint4 fetch_a(const uint16_t* __restrict__ A, int lda, int M, int m0, int k0, int kend) {
    const int t = threadIdx.x;
    if (!TA) {
        const int m = t >> 2, kq = (t & 3) * 8;
        return fetch8(A + (size_t)(m0 + m) * lda + k0 + kq, m0 + m < M, kend - (k0 + kq));
    }
    const int k = t >> 3, mq = (t & 7) * 8;
    return fetch8(A + (size_t)(k0 + k) * lda + m0 + mq, k0 + k < kend, M - (m0 + mq));
}
template <bool TA>
__device__ __forceinline__ void stage_a(const uint4& q, uint16_t (*As)[KP]) {
    const int t = threadIdx.x;
    if (!TA) {
        *reinterpret_cast<uint4*>(&As[t >> 2][(t & 3) * 8]) = q;
    } else {
        const int k = t >> 3, mq = (t & 7) * 8;
        #pragma unroll
        for (int j = 0; j < 8; ++j) As[mq + j][k] = lane16(q, j);
    }
}

// op(B)[k][n] stored as Bs[n][k]: 8 consecutive n of one k (B is K x N) or
// 8 consecutive k of one row (B is N x K)
template <bool TB>
__device__ __forceinline__ uint4 fetch_b(const uint16_t* __restrict__ B, int ldb, int N, int n0, int k0, int kend) {
    const int t = threadIdx.x;
    if (!TB) {
        const int k = t >> 3, nq = (t & 7) * 8;
        return fetch8(B + (size_t)(k0 + k) * ldb + n0 + nq, k0 + k < kend, N - (n0 + nq));
    }
    const int n = t >> 2, kq = (t & 3) * 8;
    return fetch8(B + (size_t)(n0 + n) * ldb + k0 + kq, n0 + n < N, kend - (k0 + kq));
}
template <bool TB>
__device__ __forceinline__ void stage_b(const uint4& q, uint16_t (*Bs)[KP]) {
    const int t = threadIdx.x;
    if (!TB) {
        const int k = t >> 3, nq = (t & 7) * 8;
        #pragma unroll
        for (int j = 0; j < 8; ++j) Bs[nq + j][k] = lane16(q, j);
    } else {
        *reinterpret_cast<uint4*>(&Bs[t >> 2][(t & 3) * 8]) = q;
    }
}

struct Epi {
    float* C; int ldc;
    uint16_t* Cb; int ldcb;
    const float* bias;
    int flags;
};

__device__ __forceinline__ void epilogue_store(const Epi& E, int row, int col, float x) {
    if (E.flags & ICESUR_BIAS) x = x + E.bias[col];
    if (E.flags & ICESUR_ACCUM) x = E.C[(size_t)row * E.ldc + col] + x;
    if (E.flags & ICESUR_RELU) x = fmaxf(x, 0.0f);
    if (E.C) E.C[(size_t)row * E.ldc + col] = x;
    if (E.Cb) E.Cb[(size_t)row * E.ldcb + col] = f2bf(x);
}

// grid (ceil(M/64), ceil(N/64), splits); split z covers k in [z ks, (z+1) ks)
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const uint16_t* __restrict__ A, int lda,
                                              const uint16_t* __restrict__ B, int ldb, Epi E, int ks,
                                              float* __restrict__ part, const float* __restrict__ init) {
    // two LDS buffers: the next K step's global loads are in flight while
    // the current step's MFMAs run
    __shared__ __align__(16) uint16_t As[2][TM][KP];
    __shared__ __align__(16) uint16_t Bs[2][TN][KP];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m0 = blockIdx.x * TM, n0 = blockIdx.y * TN;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    const int kbeg = blockIdx.z * ks, kend = min(K, kbeg + ks);
    f32x4 acc[2][2];
    #pragma unroll
    for (int i = 0; i < 2; ++i)
        #pragma unroll
        for (int j = 0; j < 2; ++j) {
            // accumulator column lane % 16 (rows 4 (lane / 16) + v)
            const int col = n0 + wn + j * 16 + (lane & 15);
            const float a0 = (init && col < N) ? init[col] : 0.0f;
            acc[i][j] = f32x4{a0, a0, a0, a0};
        }
    int cur = 0;
    if (kbeg < kend) {
        stage_a<TA>(fetch_a<TA>(A, lda, M, m0, kbeg, kend), As[0]);
        stage_b<TB>(fetch_b<TB>(B, ldb, N, n0, kbeg, kend), Bs[0]);
    }
    __syncthreads();
    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        const bool more = k0 + TK < kend;
        uint4 qa = make_uint4(0u, 0u, 0u, 0u), qb = qa;
        if (more) { qa = fetch_a<TA>(A, lda, M, m0, k0 + TK, kend); qb = fetch_b<TB>(B, ldb, N, n0, k0 + TK, kend); }
        bf16x8 a[2], b[2];
        #pragma unroll
        for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const bf16x8*>(&As[cur][wm + i * 16 + (lane & 15)][8 * (lane >> 4)]);
        #pragma unroll
        for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&Bs[cur][wn + j * 16 + (lane & 15)][8 * (lane >> 4)]);
        #pragma unroll
        for (int i = 0; i < 2; ++i)
            #pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        if (more) { stage_a<TA>(qa, As[cur ^ 1]); stage_b<TB>(qb, Bs[cur ^ 1]); }
        __syncthreads();
        cur ^= 1;
    }
    // D fragment: lane holds column lane % 16, rows 4 (lane / 16) + v
    #pragma unroll
    for (int i = 0; i < 2; ++i)
        #pragma unroll
        for (int j = 0; j < 2; ++j)
            #pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = m0 + wm + i * 16 + 4 * (lane >> 4) + v;
                const int col = n0 + wn + j * 16 + (lane & 15);
                if (row < M && col < N) {
                    if (part) part[((size_t)blockIdx.z * M + row) * N + col] = acc[i][j][v];
                    else epilogue_store(E, row, col, acc[i][j][v]);
                }
            }
}

// groups of 32 split partials summed in split order (the first level when
// there are more than 32 splits)
__global__ void k_splitk_group(int M, int N, int splits, const float* __restrict__ part, float* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t MN = (size_t)M * N;
    if (i >= MN) return;
    const int g = blockIdx.y, z0 = 32 * g, z1 = min(splits, z0 + 32);
    float x = part[(size_t)z0 * MN + i];
    for (int z = z0 + 1; z < z1; ++z) x = x + part[(size_t)z * MN + i];
    out[(size_t)g * MN + i] = x;
}

// sum of the split partials in split order, then the epilogue
__global__ void k_splitk_reduce(int M, int N, int splits, const float* __restrict__ part, Epi E) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)M * N) return;
    float x = part[i];
    for (int z = 1; z < splits; ++z) x = x + part[(size_t)z * M * N + i];
    epilogue_store(E, (int)(i / N), (int)(i % N), x);
}

// ── implicit-GEMM convolution ───────────────────────────────────────────────
// 3x3, stride 2, padding 1, NHWC: out[m][n] = epi(sum_k A[m][k] Wt[n][k]) with
// m = (b, ho, wo) and k = tap C + c; the A tile is gathered from X while it is
// staged into LDS, so no im2col matrix is written.  128 x TN tile per
// workgroup of 4 waves (each 32 rows x TN as 2 x TN/16 MFMA tiles), K in
// steps of 32, two LDS buffers: the next step's global loads are in flight
// while the current step's MFMAs run.  Gather modes: CONV_VEC (C a multiple
// of 32: a 32-wide k step lies inside one tap, 16-byte chunks), CONV_PAIR
// (C = 2: one dword per tap, 9 taps = 18 k in one step), CONV_GEN (element
// by element).  Wt is the weight matrix transposed, (N, K) row-major.  The
// output tile goes through LDS so that global stores are 16-byte rows.
constexpr int CM = 128, CK = 32, CKP = CK + 8;
constexpr int CONV_GEN = 0, CONV_VEC = 1, CONV_PAIR = 2;

template <int TN, int MODE>
__global__ __launch_bounds__(256) void k_conv(const uint16_t* __restrict__ X, int B, int H, int W, int C, int Ho,
                                              int Wo, const uint16_t* __restrict__ Wt, int N,
                                              const float* __restrict__ bias, uint16_t* __restrict__ out, int flags) {
    constexpr int NBUF = MODE == CONV_PAIR ? 1 : 2;           // C = 2: K = 18 is one step
    constexpr int A_BYTES = NBUF * CM * CKP * 2, B_BYTES = NBUF * TN * CKP * 2;
    constexpr int O_LD = TN + 8;
    constexpr int O_BYTES = CM * O_LD * 2;
    constexpr int L_BYTES = (A_BYTES + B_BYTES) > O_BYTES ? (A_BYTES + B_BYTES) : O_BYTES;
    __shared__ __align__(16) uint8_t lds[L_BYTES];
    uint16_t (*As)[CM][CKP] = reinterpret_cast<uint16_t (*)[CM][CKP]>(lds);
    uint16_t (*Bs)[TN][CKP] = reinterpret_cast<uint16_t (*)[TN][CKP]>(lds + A_BYTES);
    uint16_t (*Os)[O_LD] = reinterpret_cast<uint16_t (*)[O_LD]>(lds);
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int M = B * Ho * Wo, K = 9 * C;
    const int m0 = blockIdx.x * CM, n0 = blockIdx.y * TN;
    // this thread's rows: VEC t/4 and t/4 + 64 (8-channel chunk t % 4); PAIR t/2 (taps 0-4 or 5-8)
    int rb[2] = {0, 0}, rh[2] = {0, 0}, rw[2] = {0, 0};
    bool rv[2] = {false, false};
    const int q = t & 3;
    #pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int m = MODE == CONV_VEC ? m0 + (t >> 2) + 64 * r : m0 + (t >> 1);
        if (MODE == CONV_GEN || (MODE == CONV_PAIR && r > 0)) continue;
        rv[r] = m < M;
        const int mm = rv[r] ? m : 0;
        rb[r] = mm / (Ho * Wo);
        const int rem = mm - rb[r] * Ho * Wo;
        rh[r] = 2 * (rem / Wo) - 1;
        rw[r] = 2 * (rem % Wo) - 1;
    }
    const int p0 = (t & 1) ? 5 : 0, pn = (t & 1) ? 4 : 5;   // PAIR: first tap, tap count
    constexpr int BCH = TN * CK / 8;                        // 16-byte chunks of the B tile
    uint4 pa[2], pb[(BCH + 255) / 256];
    uint32_t pp[5];
    uint16_t ga[16];
    auto fetch = [&](int k0) {
        if (MODE == CONV_VEC) {
            const int tap = k0 / C, c0 = k0 - tap * C + 8 * q;
            const int kh = tap / 3, kw = tap - 3 * (tap / 3);
            #pragma unroll
            for (int r = 0; r < 2; ++r) {
                const int h = rh[r] + kh, ww = rw[r] + kw;
                if (rv[r] && h >= 0 && h < H && ww >= 0 && ww < W)
                    pa[r] = *reinterpret_cast<const uint4*>(X + (((size_t)rb[r] * H + h) * W + ww) * C + c0);
                else
                    pa[r] = make_uint4(0u, 0u, 0u, 0u);
            }
        } else if (MODE == CONV_PAIR) {
            #pragma unroll
            for (int i = 0; i < 5; ++i) {
                const int tap = p0 + i;
                const int h = rh[0] + tap / 3, ww = rw[0] + tap % 3;
                pp[i] = (i < pn && rv[0] && h >= 0 && h < H && ww >= 0 && ww < W)
                            ? *reinterpret_cast<const uint32_t*>(X + (((size_t)rb[0] * H + h) * W + ww) * 2)
                            : 0u;
            }
        } else {
            #pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int e = t + 256 * i, row = e >> 5, k = k0 + (e & 31);
                const int m = m0 + row;
                uint16_t v = 0;
                if (m < M && k < K) {
                    const int b = m / (Ho * Wo), rem = m - b * Ho * Wo;
                    const int tap = k / C, c = k - tap * C;
                    const int h = 2 * (rem / Wo) - 1 + tap / 3, ww = 2 * (rem % Wo) - 1 + tap % 3;
                    if (h >= 0 && h < H && ww >= 0 && ww < W) v = X[(((size_t)b * H + h) * W + ww) * C + c];
                }
                ga[i] = v;
            }
        }
        #pragma unroll
        for (int i = 0; i < (BCH + 255) / 256; ++i) {
            const int ch = t + 256 * i;
            if (ch < BCH) {
                const int n = ch >> 2, kq = k0 + 8 * (ch & 3);
                if (MODE == CONV_VEC) {
                    pb[i] = (n0 + n < N) ? *reinterpret_cast<const uint4*>(Wt + (size_t)(n0 + n) * K + kq)
                                         : make_uint4(0u, 0u, 0u, 0u);
                } else {
                    uint16_t v[8];
                    #pragma unroll
                    for (int j = 0; j < 8; ++j)
                        v[j] = (n0 + n < N && kq + j < K) ? Wt[(size_t)(n0 + n) * K + kq + j] : (uint16_t)0;
                    pb[i] = make_uint4(v[0] | ((uint32_t)v[1] << 16), v[2] | ((uint32_t)v[3] << 16),
                                       v[4] | ((uint32_t)v[5] << 16), v[6] | ((uint32_t)v[7] << 16));
                }
            }
        }
    };
    auto stage = [&](int buf) {
        if (MODE == CONV_VEC) {
            #pragma unroll
            for (int r = 0; r < 2; ++r) *reinterpret_cast<uint4*>(&As[buf][(t >> 2) + 64 * r][8 * q]) = pa[r];
        } else if (MODE == CONV_PAIR) {
            uint16_t* row = As[buf][t >> 1];
            #pragma unroll
            for (int i = 0; i < 5; ++i)
                if (i < pn) *reinterpret_cast<uint32_t*>(row + 2 * (p0 + i)) = pp[i];
            if (t & 1) {                                      // k = 18 .. 31 are zero
                #pragma unroll
                for (int c = 18; c < 32; c += 2) *reinterpret_cast<uint32_t*>(row + c) = 0u;
            }
        } else {
            #pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int e = t + 256 * i;
                As[buf][e >> 5][e & 31] = ga[i];
            }
        }
        #pragma unroll
        for (int i = 0; i < (BCH + 255) / 256; ++i) {
            const int ch = t + 256 * i;
            if (ch < BCH) *reinterpret_cast<uint4*>(&Bs[buf][ch >> 2][8 * (ch & 3)]) = pb[i];
        }
    };
    f32x4 acc[2][TN / 16];
    #pragma unroll
    for (int i = 0; i < 2; ++i)
        #pragma unroll
        for (int j = 0; j < TN / 16; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    fetch(0);
    stage(0);
    __syncthreads();
    int cur = 0;
    for (int k0 = 0; k0 < K; k0 += CK) {
        const bool more = k0 + CK < K;
        if (more) fetch(k0 + CK);
        bf16x8 a[2];
        #pragma unroll
        for (int i = 0; i < 2; ++i)
            a[i] = *reinterpret_cast<const bf16x8*>(&As[cur][w * 32 + i * 16 + (lane & 15)][8 * (lane >> 4)]);
        #pragma unroll
        for (int j = 0; j < TN / 16; ++j) {
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[cur][j * 16 + (lane & 15)][8 * (lane >> 4)]);
            #pragma unroll
            for (int i = 0; i < 2; ++i) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b, acc[i][j], 0, 0, 0);
        }
        if (more) stage(cur ^ 1);
        __syncthreads();
        if (NBUF == 2) cur ^= 1;
    }
    // epilogue into LDS (the loop's last barrier retired every read of As / Bs)
    #pragma unroll
    for (int j = 0; j < TN / 16; ++j) {
        const int cl = j * 16 + (lane & 15);
        const float bc = (n0 + cl < N && (flags & ICESUR_BIAS)) ? bias[n0 + cl] : 0.0f;
        #pragma unroll
        for (int i = 0; i < 2; ++i)
            #pragma unroll
            for (int v = 0; v < 4; ++v) {
                float x = acc[i][j][v] + bc;
                if (flags & ICESUR_RELU) x = fmaxf(x, 0.0f);
                Os[w * 32 + i * 16 + 4 * (lane >> 4) + v][cl] = f2bf(x);
            }
    }
    __syncthreads();
    const bool vo = (N & 7) == 0 && (((uintptr_t)out) & 15) == 0;
    constexpr int RCH = TN / 8;                        // 16-byte chunks per tile row
    for (int ch = t; ch < CM * RCH; ch += 256) {
        const int r = ch / RCH, c8 = 8 * (ch % RCH);
        const int row = m0 + r, col = n0 + c8;
        if (row >= M || col >= N) continue;
        if (vo) {
            *reinterpret_cast<uint4*>(out + (size_t)row * N + col) = *reinterpret_cast<const uint4*>(&Os[r][c8]);
        } else {
            for (int j = 0; j < 8 && col + j < N; ++j) out[(size_t)row * N + col + j] = Os[r][c8 + j];
        }
    }
}

// ── the first two convolutions fused (inference) ────────────────────────────
// conv 1 (64x64x2 -> 32x32x32) and conv 2 (-> 16x16x64), 3x3 stride 2
// padding 1, for R2 conv-2 output rows [R2 h, R2 h + R2) of one sample per
// workgroup: the 2 R2 + 1 conv-1 rows those read are computed into LDS -- one
// MFMA per 16 pixels x 16 channels with K = 18 padded to 32, the operands of
// k_conv's C = 2 path -- and conv 2 reads them from LDS tap by tap, k_conv's
// K order, so the output is bit-identical to the two k_conv launches while
// a sample's 64 KB conv-1 activation never leaves the CU.  Conv-1 pixels are
// held by column parity (stride-2 reads of neighbouring output pixels hit
// neighbouring 80-byte LDS rows).  Wt1 (32, 18) and Wt2 (64, 288) are the
// transposed weights.
constexpr int F1_CP = 40, F2_LD = 64 + 8;

template <int R2>
__global__ __launch_bounds__(256) void k_conv12(const uint16_t* __restrict__ X, const uint16_t* __restrict__ Wt1,
                                                const float* __restrict__ b1, const uint16_t* __restrict__ Wt2,
                                                const float* __restrict__ b2, int flags, uint16_t* __restrict__ out) {
    constexpr int F1_ROWS = 2 * R2 + 1, FX_ROWS = 2 * F1_ROWS + 1;
    constexpr int XS_BYTES = FX_ROWS * 64 * 4;
    constexpr int W1_BYTES = 32 * KP * 2;
    constexpr int O1_BYTES = F1_ROWS * 2 * 16 * F1_CP * 2;
    constexpr int OS_BYTES = 16 * R2 * F2_LD * 2;
    constexpr int L1_BYTES = XS_BYTES + W1_BYTES + O1_BYTES;
    constexpr int L_BYTES = L1_BYTES > OS_BYTES ? L1_BYTES : OS_BYTES;
    __shared__ __align__(16) uint8_t lds[L_BYTES];
    uint32_t (*Xs)[64] = reinterpret_cast<uint32_t (*)[64]>(lds);                 // 2 channels per dword
    uint16_t (*W1s)[KP] = reinterpret_cast<uint16_t (*)[KP]>(lds + XS_BYTES);
    uint16_t (*O1)[2][16][F1_CP] = reinterpret_cast<uint16_t (*)[2][16][F1_CP]>(lds + XS_BYTES + W1_BYTES);
    uint16_t (*Os)[F2_LD] = reinterpret_cast<uint16_t (*)[F2_LD]>(lds);           // after conv 2's reads
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, kc = lane >> 4;
    const int b = blockIdx.x, h = blockIdx.y;
    // conv 2's weight fragments of this lane (output channel 16 w + lane % 16,
    // k chunk kc of each tap), issued first so they arrive during conv 1
    const int n2 = 16 * w + (lane & 15), w2 = lane & 15;
    bf16x8 bw[9];
    #pragma unroll
    for (int tap = 0; tap < 9; ++tap)
        bw[tap] = *reinterpret_cast<const bf16x8*>(Wt2 + (size_t)n2 * 288 + 32 * tap + 8 * kc);
    // input rows 4 R2 h - 3 .. 4 R2 h + 4 R2 - 1 (zeros outside the image), 4 pixels a chunk
    const uint32_t* Xb = reinterpret_cast<const uint32_t*>(X) + (size_t)b * 64 * 64;
    for (int ch = t; ch < FX_ROWS * 16; ch += 256) {
        const int i = ch >> 4, c4 = (ch & 15) * 4, x = 4 * R2 * h - 3 + i;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (x >= 0 && x < 64) v = *reinterpret_cast<const uint4*>(Xb + x * 64 + c4);
        *reinterpret_cast<uint4*>(&Xs[i][c4]) = v;
    }
    for (int e = t; e < 32 * 32; e += 256) {
        const int n = e >> 5, k = e & 31;
        W1s[n][k] = k < 18 ? Wt1[n * 18 + k] : (uint16_t)0;
    }
    __syncthreads();
    // conv 1: tile q = (local row j, half hh), pixels w1 = 16 hh + lane % 16;
    // this lane's k chunk holds taps 4 kc .. 4 kc + 3 (k = 2 tap + channel)
    for (int q = w; q < 2 * F1_ROWS; q += 4) {
        const int j = q >> 1, hh = q & 1;
        const int r1 = 2 * R2 * h - 1 + j;             // -1: conv 2's zero padding
        const int w1 = 16 * hh + (lane & 15);
        uint4 d = make_uint4(0u, 0u, 0u, 0u);
        uint32_t* dv = reinterpret_cast<uint32_t*>(&d);
        #pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int tap = 4 * kc + i;
            if (tap < 9) {
                const int kh = tap / 3, kw = tap - 3 * (tap / 3);
                const int col = 2 * w1 - 1 + kw;
                if (col >= 0 && col < 64) dv[i] = Xs[2 * j + kh][col];
            }
        }
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(&d);
        #pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const bf16x8 bb = *reinterpret_cast<const bf16x8*>(&W1s[nt * 16 + (lane & 15)][8 * kc]);
            const f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, f32x4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
            const int n = nt * 16 + (lane & 15);
            const float bc = (flags & ICESUR_BIAS) ? b1[n] : 0.0f;
            #pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int px = 16 * hh + 4 * kc + v;
                float x = acc[v] + bc;
                if (flags & ICESUR_RELU) x = fmaxf(x, 0.0f);
                O1[j][px & 1][px >> 1][n] = r1 >= 0 ? f2bf(x) : (uint16_t)0;
            }
        }
    }
    __syncthreads();
    // conv 2: wave w owns output channels 16 w .. 16 w + 15 of the R2 rows
    f32x4 acc2[R2];
    #pragma unroll
    for (int rr = 0; rr < R2; ++rr) acc2[rr] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    #pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
        const int kh = tap / 3, kw = tap - 3 * (tap / 3);
        const bf16x8 bb = bw[tap];
        const int par = kw == 1 ? 0 : 1, u = kw == 0 ? w2 - 1 : w2;   // conv-1 column 2 w2 - 1 + kw
        #pragma unroll
        for (int rr = 0; rr < R2; ++rr) {
            bf16x8 a = bf16x8{};
            if (u >= 0) a = *reinterpret_cast<const bf16x8*>(&O1[2 * rr + kh][par][u][8 * kc]);
            acc2[rr] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc2[rr], 0, 0, 0);
        }
    }
    __syncthreads();   // every wave's conv-1 reads are done before Os reuses the space
    const float bc2 = (flags & ICESUR_BIAS) ? b2[n2] : 0.0f;
    #pragma unroll
    for (int rr = 0; rr < R2; ++rr)
        #pragma unroll
        for (int v = 0; v < 4; ++v) {
            float x = acc2[rr][v] + bc2;
            if (flags & ICESUR_RELU) x = fmaxf(x, 0.0f);
            Os[rr * 16 + 4 * kc + v][n2] = f2bf(x);
        }
    __syncthreads();
    // the R2 x 16 x 64 output block is contiguous in NHWC
    uint16_t* ob = out + ((size_t)b * 16 + R2 * h) * 16 * 64;
    for (int ch = t; ch < 16 * R2 * 8; ch += 256) {
        const int P = ch >> 3, c8 = (ch & 7) * 8;
        *reinterpret_cast<uint4*>(ob + (size_t)P * 64 + c8) = *reinterpret_cast<const uint4*>(&Os[P][c8]);
    }
}

// (K, N) -> (N, K) bf16
__global__ void k_transpose16(const uint16_t* __restrict__ a, int K, int N, uint16_t* __restrict__ at) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)K * N) return;
    const int n = (int)(i / K), k = (int)(i % K);
    at[i] = a[(size_t)k * N + n];
}

__global__ void k_im2col(const uint16_t* __restrict__ X, int B, int H, int W, int C, int Ho, int Wo,
                         uint16_t* __restrict__ col) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)B * Ho * Wo * 9 * C;
    if (idx >= total) return;
    const int c = (int)(idx % C);
    const size_t r1 = idx / C;
    const int tap = (int)(r1 % 9);
    const size_t row = r1 / 9;
    const int wo = (int)(row % Wo), ho = (int)((row / Wo) % Ho), b = (int)(row / ((size_t)Wo * Ho));
    const int h = 2 * ho + tap / 3 - 1, w = 2 * wo + tap % 3 - 1;
    col[idx] = (h >= 0 && h < H && w >= 0 && w < W) ? X[(((size_t)b * H + h) * W + w) * C + c] : (uint16_t)0;
}

__global__ void k_col2im(const float* __restrict__ dcol, int B, int H, int W, int C, int Ho, int Wo,
                         float* __restrict__ dx) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)B * H * W * C) return;
    const int c = (int)(idx % C);
    const size_t p = idx / C;
    const int w = (int)(p % W), h = (int)((p / W) % H), b = (int)(p / ((size_t)W * H));
    float s = 0.0f;
    for (int kh = 0; kh < 3; ++kh) {
        const int hh = h + 1 - kh;
        if (hh < 0 || (hh & 1) || (hh >> 1) >= Ho) continue;
        for (int kw = 0; kw < 3; ++kw) {
            const int ww = w + 1 - kw;
            if (ww < 0 || (ww & 1) || (ww >> 1) >= Wo) continue;
            const size_t row = ((size_t)b * Ho + (hh >> 1)) * Wo + (ww >> 1);
            s = s + dcol[row * 9 * C + (size_t)(kh * 3 + kw) * C + c];
        }
    }
    dx[idx] = s;
}

__global__ void k_relu_back(float* __restrict__ g, int ldg, const uint16_t* __restrict__ y, int ldy, int rows,
                            int cols, uint16_t* __restrict__ gb, int ldgb) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)rows * cols) return;
    const int r = (int)(idx / cols), c = (int)(idx % cols);
    float v = g[(size_t)r * ldg + c];
    v = (bf2f(y[(size_t)r * ldy + c]) > 0.0f) ? v : 0.0f;
    g[(size_t)r * ldg + c] = v;
    if (gb) gb[(size_t)r * ldgb + c] = f2bf(v);
}

// block (b, cb) of 256 threads = 64 columns x 4 row phases: phase p sums rows
// 256 b + p, + 4, + 8, ... in order; work[b * cols + c] = (s0 + s1) + (s2 + s3)
__global__ __launch_bounds__(256) void k_colsum_part(const float* __restrict__ x, int ldx, int rows, int cols,
                                                     float* __restrict__ work) {
    __shared__ float sp[4][64];
    const int cl = threadIdx.x & 63, p = threadIdx.x >> 6;
    const int c = blockIdx.y * 64 + cl;
    const int r0 = blockIdx.x * 256, r1 = min(rows, r0 + 256);
    float s = 0.0f;
    if (c < cols)
        for (int r = r0 + p; r < r1; r += 4) s = s + x[(size_t)r * ldx + c];
    sp[p][cl] = s;
    __syncthreads();
    if (p == 0 && c < cols) work[(size_t)blockIdx.x * cols + c] = (sp[0][cl] + sp[1][cl]) + (sp[2][cl] + sp[3][cl]);
}

// block cb of 256 threads = 16 columns x 16 phases: phase g sums the block sums
// g, g + 16, ... in order; the 16 phase sums are added as a pairwise tree
__global__ __launch_bounds__(256) void k_colsum_final(const float* __restrict__ work, int nblk, int cols,
                                                      float* __restrict__ out, int accumulate) {
    __shared__ float sp[16][16];
    const int cl = threadIdx.x & 15, g = threadIdx.x >> 4;
    const int c = blockIdx.x * 16 + cl;
    float s = 0.0f;
    if (c < cols)
        for (int b = g; b < nblk; b += 16) s = s + work[(size_t)b * cols + c];
    sp[g][cl] = s;
    __syncthreads();
    if (g == 0 && c < cols) {
        float v[16];
        #pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = sp[i][cl];
        #pragma unroll
        for (int w = 1; w < 16; w *= 2)
            #pragma unroll
            for (int i = 0; i + w < 16; i += 2 * w) v[i] = v[i] + v[i + w];
        out[c] = accumulate ? out[c] + v[0] : v[0];
    }
}

__global__ void k_to_bf16(const float* __restrict__ x, int64_t n, uint16_t* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f2bf(x[i]);
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float c1, float c2,
                       float wd, uint16_t* __restrict__ wb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * (gi * gi);
    m[i] = mi; v[i] = vi;
    const float mh = mi / c1, vh = vi / c2;
    const float w0 = w[i];
    const float wi = w0 - lr * (mh / (sqrtf(vh) + eps) + wd * w0);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

// one workgroup: per-row sums in parallel, rows added in order by thread 0
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ y, const float* __restrict__ t,
                                             const float* __restrict__ wt, int B, int D, float* __restrict__ dy,
                                             uint16_t* __restrict__ dyb, float* __restrict__ loss) {
    __shared__ float rs[256];
    const float inv = 1.0f / (float)((int64_t)B * D);
    float total = 0.0f;
    for (int r0 = 0; r0 < B; r0 += 256) {
        const int r = r0 + threadIdx.x;
        float s = 0.0f;
        if (r < B) {
            for (int o = 0; o < D; ++o) {
                const size_t i = (size_t)r * D + o;
                const float d = y[i] - t[i];
                s = s + wt[o] * (d * d);
                const float gi = 2.0f * wt[o] * d * inv;
                dy[i] = gi;
                if (dyb) dyb[i] = f2bf(gi);
            }
        }
        rs[threadIdx.x] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            const int m = min(256, B - r0);
            for (int q = 0; q < m; ++q) total = total + rs[q];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[0] = total * inv;
}

__global__ void k_concat(const uint16_t* __restrict__ feat, int ldf, int F, const float* __restrict__ extra, int lde,
                         int E, int B, uint16_t* __restrict__ z, int ldz) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)B * ldz) return;
    const int b = (int)(idx / ldz), j = (int)(idx % ldz);
    uint16_t v = 0;
    if (j < F) v = feat[(size_t)b * ldf + j];
    else if (j < F + E) v = f2bf(extra[(size_t)b * lde + (j - F)]);
    z[idx] = v;
}

__global__ void k_split(const float* __restrict__ dz, int lddz, int F, int B, float* __restrict__ dfeat) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)B * F) return;
    dfeat[idx] = dz[(idx / F) * lddz + idx % F];
}

// thread per (pose, element): splat the element's equivalent disc onto the
// cells whose centres it covers (integer atomics: the sum is order-free)
__global__ void k_raster_splat(const float* __restrict__ pos, const float* __restrict__ area,
                               const float* __restrict__ thick, int n, const float* __restrict__ pose, int B, int G,
                               float cell, float x_lo, int32_t* __restrict__ acc) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)B * n) return;
    const int b = (int)(idx / n), e = (int)(idx % n);
    const float px = pose[3 * b], py = pose[3 * b + 1], ph = pose[3 * b + 2];
    const float c = cosf(ph), s = sinf(ph);
    const float dx = pos[e] - px, dy = pos[n + e] - py;
    const float u = c * dx + s * dy, v = -s * dx + c * dy;      // ship frame
    const float r = sqrtf(area[e] / 3.14159265f);
    const float y_lo = -0.5f * (float)G * cell;
    const int i0 = max(0, (int)floorf((u - r - x_lo) / cell)), i1 = min(G - 1, (int)floorf((u + r - x_lo) / cell));
    const int j0 = max(0, (int)floorf((v - r - y_lo) / cell)), j1 = min(G - 1, (int)floorf((v + r - y_lo) / cell));
    const int ca = (int)(cell * cell * 4096.0f);
    const int ch = (int)(cell * cell * thick[e] * 4096.0f);
    for (int i = i0; i <= i1; ++i) {
        const float cu = x_lo + ((float)i + 0.5f) * cell;
        for (int j = j0; j <= j1; ++j) {
            const float cv = y_lo + ((float)j + 0.5f) * cell;
            if ((cu - u) * (cu - u) + (cv - v) * (cv - v) > r * r) continue;
            int32_t* a = acc + (((size_t)b * G + i) * G + j) * 2;
            atomicAdd(a, ca);
            atomicAdd(a + 1, ch);
        }
    }
}

// Convex polygons (world vertices (n, maxV) float2, nv[e] used) sampled at
// SS x SS points per cell of the ship-frame window; one thread per (pose,
// element); integer atomics, so the raster does not depend on thread order.
constexpr int RASTER_SS = 4;

__global__ void k_raster_poly(const float2* __restrict__ verts, const int32_t* __restrict__ nv,
                              const float* __restrict__ thick, int n, int maxV, const float* __restrict__ pose,
                              int B, int G, float cell, float x_lo, int32_t* __restrict__ acc) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)B * n) return;
    const int b = (int)(idx / n), e = (int)(idx % n);
    const int m = min(nv[e], maxV);
    if (m < 3) return;
    const float px = pose[3 * b], py = pose[3 * b + 1], ph = pose[3 * b + 2];
    const float c = cosf(ph), s = sinf(ph);
    const float2* P = verts + (size_t)e * maxV;
    float u_lo = INFINITY, u_hi = -INFINITY, v_lo = INFINITY, v_hi = -INFINITY, area2 = 0.0f;
    float pu = 0.0f, pv = 0.0f;
    for (int k = 0; k <= m; ++k) {
        const float2 w = P[k % m];
        const float dx = w.x - px, dy = w.y - py;
        const float u = c * dx + s * dy, v = -s * dx + c * dy;
        if (k > 0) area2 += pu * v - pv * u;
        u_lo = fminf(u_lo, u); u_hi = fmaxf(u_hi, u); v_lo = fminf(v_lo, v); v_hi = fmaxf(v_hi, v);
        pu = u; pv = v;
    }
    const float orient = area2 >= 0.0f ? 1.0f : -1.0f;
    const float y_lo = -0.5f * (float)G * cell;
    const int i0 = max(0, (int)floorf((u_lo - x_lo) / cell)), i1 = min(G - 1, (int)floorf((u_hi - x_lo) / cell));
    const int j0 = max(0, (int)floorf((v_lo - y_lo) / cell)), j1 = min(G - 1, (int)floorf((v_hi - y_lo) / cell));
    if (i0 > i1 || j0 > j1) return;
    const int w1 = (int)(cell * cell * 4096.0f / (float)(RASTER_SS * RASTER_SS));
    const int wh = (int)(cell * cell * thick[e] * 4096.0f / (float)(RASTER_SS * RASTER_SS));
    const float h = cell / (float)RASTER_SS;
    for (int i = i0; i <= i1; ++i) {
        for (int j = j0; j <= j1; ++j) {
            int cnt = 0;
            for (int a = 0; a < RASTER_SS; ++a) {
                const float qu = x_lo + (float)i * cell + ((float)a + 0.5f) * h;
                for (int q = 0; q < RASTER_SS; ++q) {
                    const float qv = y_lo + (float)j * cell + ((float)q + 0.5f) * h;
                    bool in = true;
                    float au = 0.0f, av = 0.0f;
                    for (int k = 0; k <= m && in; ++k) {
                        const float2 w = P[k % m];
                        const float dx = w.x - px, dy = w.y - py;
                        const float u = c * dx + s * dy, v = -s * dx + c * dy;
                        if (k > 0) in = orient * ((u - au) * (qv - av) - (v - av) * (qu - au)) >= 0.0f;
                        au = u; av = v;
                    }
                    cnt += in ? 1 : 0;
                }
            }
            if (cnt) {
                int32_t* d = acc + (((size_t)b * G + i) * G + j) * 2;
                atomicAdd(d, cnt * w1);
                atomicAdd(d + 1, cnt * wh);
            }
        }
    }
}

__global__ void k_raster_out(int32_t* __restrict__ acc, int B, int G, float cell, uint16_t* __restrict__ out) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)B * G * G) return;
    const float a = (float)acc[2 * idx], ah = (float)acc[2 * idx + 1];
    const float ca = cell * cell * 4096.0f;
    out[2 * idx] = f2bf(fminf(a / ca, 1.0f));
    out[2 * idx + 1] = f2bf(a > 0.0f ? ah / a : 0.0f);
    acc[2 * idx] = 0;
    acc[2 * idx + 1] = 0;
}

// ── KE costmap (jax_costmap.py:98-141, 177-330) ─────────────────────────────
// Per element the AUTO-IceNav KE constant ke = w v^2 m M (m + 2 M) / (2 (M +
// m)^2) s (binary64, clipped to [0, 1e8]); per cell the maximum over elements
// of ke max(0, (r_e^2 - d^2) / r_e^2), r_e = r (1 + b) + res / 2, by integer
// atomic max of the non-negative float bits (order-free).
__global__ void k_ke_zero(int32_t* __restrict__ bits, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) bits[i] = 0;
}

__global__ void k_ke_splat(const float* __restrict__ pos, const float* __restrict__ radius,
                           const float* __restrict__ mass, int n, int rows, int cols, float x0, float y0, float res,
                           double w, double v, double Ms, double scale, float buf, int32_t* __restrict__ bits) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double m = mass[e];
    double ke = w * v * v * m * Ms * (m + 2.0 * Ms) / (2.0 * (Ms + m) * (Ms + m)) * scale;
    ke = fmin(fmax(ke, 0.0), 1e8);
    const float kef = (float)ke;
    const float px = pos[e], py = pos[n + e];
    const float er = radius[e] * (1.0f + buf) + 0.5f * res;
    const float er2 = er * er;
    const int c0 = max(0, (int)floorf((px - er - x0) / res)), c1 = min(cols - 1, (int)floorf((px + er - x0) / res));
    const int r0 = max(0, (int)floorf((py - er - y0) / res)), r1 = min(rows - 1, (int)floorf((py + er - y0) / res));
    for (int r = r0; r <= r1; ++r) {
        const float cy = y0 + ((float)r + 0.5f) * res;
        for (int c = c0; c <= c1; ++c) {
            const float cx = x0 + ((float)c + 0.5f) * res;
            const float d2 = (cx - px) * (cx - px) + (cy - py) * (cy - py);
            const float val = kef * fmaxf(0.0f, (er2 - d2) / er2);
            if (val > 0.0f) atomicMax(bits + (size_t)r * cols + c, __float_as_int(val));
        }
    }
}

// open-water floor, then the maximum over a (2 half + 1)^2 window with edge
// padding (the reference's reduce_window over the full square window)
__global__ void k_ke_dilate(const int32_t* __restrict__ bits, int rows, int cols, float base, int half,
                            float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * cols) return;
    const int r = i / cols, c = i % cols;
    float m = base;
    for (int dr = -half; dr <= half; ++dr) {
        const int rr = min(max(r + dr, 0), rows - 1);
        for (int dc = -half; dc <= half; ++dc) {
            const int cc = min(max(c + dc, 0), cols - 1);
            m = fmaxf(m, __int_as_float(bits[(size_t)rr * cols + cc]));
        }
    }
    out[i] = m;
}

// ── MPPI update (run_mppi, jax_mppi_planner.py:34-115) ──────────────────────
// beta = min c, w_i = exp(-(c_i - beta) / T), v = clip(sum_i w_i u_i / sum_i w_i,
// -1, 1).  Fixed-order sums: 256 interleaved running sums added as a pairwise
// tree (weights), samples in chunks of 256 (each chunk in order, the chunks in
// order) for the weighted sum.
__device__ __forceinline__ float lds_tree256(float* sh, float x, bool take_min) {
    sh[threadIdx.x] = x;
    __syncthreads();
    for (int w = 1; w < 256; w *= 2) {
        if ((threadIdx.x & (2 * w - 1)) == 0)
            sh[threadIdx.x] = take_min ? fminf(sh[threadIdx.x], sh[threadIdx.x + w]) : sh[threadIdx.x] + sh[threadIdx.x + w];
        __syncthreads();
    }
    const float r = sh[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(256) void k_mppi_stats(const float* __restrict__ c, int S, float T, float* __restrict__ st) {
    __shared__ float sh[256];
    float m = INFINITY;
    for (int i = threadIdx.x; i < S; i += 256) m = fminf(m, c[i]);
    const float beta = lds_tree256(sh, m, true);
    float z = 0.0f;
    for (int i = threadIdx.x; i < S; i += 256) z = z + expf(-(c[i] - beta) / T);
    const float eta = lds_tree256(sh, z, false);
    if (threadIdx.x == 0) { st[0] = beta; st[1] = eta; }
}

// block b: part[b][d] = sum over samples 256 b .. 256 b + 255 of w_i u[i][d], in order
__global__ __launch_bounds__(256) void k_mppi_part(const float* __restrict__ c, const float* __restrict__ u, int S, int D,
                                                   float T, const float* __restrict__ st, float* __restrict__ part) {
    __shared__ float w[256];
    const int i0 = blockIdx.x * 256, n = min(256, S - i0);
    const float beta = st[0];
    if ((int)threadIdx.x < n) w[threadIdx.x] = expf(-(c[i0 + threadIdx.x] - beta) / T);
    __syncthreads();
    for (int d = threadIdx.x; d < D; d += 256) {
        float a = 0.0f;
        for (int i = 0; i < n; ++i) a = a + w[i] * u[(size_t)(i0 + i) * D + d];
        part[(size_t)blockIdx.x * D + d] = a;
    }
}

__global__ void k_mppi_final(const float* __restrict__ part, int nb, int D, const float* __restrict__ st,
                             float* __restrict__ out) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= D) return;
    float a = 0.0f;
    for (int b = 0; b < nb; ++b) a = a + part[(size_t)b * D + d];
    out[d] = fminf(fmaxf(a / st[1], -1.0f), 1.0f);
}

// ── MBD reverse-diffusion update (run_mbd, jax_mbd_planner.py:1363-1396) ────
// The weighted mean Ybar_new = sum_i w_i Y0s_i / sum_i w_i has run_mppi's
// weights (softmax of -(c - min c) / T); then the score step of the
// diffusion, per dimension d:
//   Yi = Ybar sqrt(ab_i); score = (-Yi + sqrt(ab_i) Ybar_new) / (1 - ab_i)
//   Yim1 = (Yi + (1 - ab_i) score) / sqrt(a_i); out = clip(Yim1 / sqrt(ab_prev), -1, 1)
__global__ void k_mbd_final(const float* __restrict__ part, int nb, int D, const float* __restrict__ st,
                            const float* ybar, float ab, float a, float ab_prev, float* out) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= D) return;
    float acc = 0.0f;
    for (int b = 0; b < nb; ++b) acc = acc + part[(size_t)b * D + d];
    const float ynew = acc / st[1];
    const float sab = sqrtf(ab);
    const float yi = ybar[d] * sab;
    const float score = (1.0f / (1.0f - ab)) * ((-yi) + sab * ynew);
    const float yim1 = (1.0f / sqrtf(a)) * (yi + (1.0f - ab) * score);
    out[d] = fminf(fmaxf(yim1 / sqrtf(ab_prev), -1.0f), 1.0f);
}

}  // namespace icesur

using namespace icesur;

#define SUR_CHECK(cond, msg) do { if (!(cond)) { icedem_set_error(msg); return ICEDEM_E_INVALID; } } while (0)
#define SUR_LAUNCHED() do { hipError_t _e = hipGetLastError(); if (_e != hipSuccess) { icedem_set_error(std::string("launch: ") + hipGetErrorString(_e)); return ICEDEM_E_HIP; } } while (0)

static unsigned grid1(size_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

extern "C" {

int icesur_gemm(int32_t trans_a, int32_t trans_b, int32_t M, int32_t N, int32_t K, const uint16_t* d_a, int32_t lda,
                const uint16_t* d_b, int32_t ldb, float* d_c, int32_t ldc, uint16_t* d_cb, int32_t ldcb,
                const float* d_bias, int32_t flags, int32_t split_k, float* d_work, void* stream) {
    SUR_CHECK(M >= 0 && N >= 0 && K >= 0 && d_a && d_b && (d_c || d_cb), "icesur_gemm: bad arguments");
    SUR_CHECK(!(flags & ICESUR_ACCUM) || d_c, "icesur_gemm: accumulate needs fp32 C");
    SUR_CHECK(!(flags & ICESUR_BIAS) || d_bias, "icesur_gemm: bias missing");
    SUR_CHECK(split_k >= 1 && (split_k == 1 || d_work), "icesur_gemm: split_k needs a workspace");
    SUR_CHECK(!(flags & ICESUR_ROWINIT) || (split_k == 1 && d_work), "icesur_gemm: row start needs split_k 1 and d_work");
    SUR_CHECK((trans_a ? lda >= M : lda >= K) && (trans_b ? ldb >= K : ldb >= N), "icesur_gemm: leading dimension");
    if (M == 0 || N == 0) return ICEDEM_OK;
    hipStream_t s = (hipStream_t)stream;
    Epi E{d_c, ldc, d_cb, ldcb, d_bias, flags};
    int ks = K;
    int splits = 1;
    if (split_k > 1) {
        ks = ((K + split_k - 1) / split_k + TK - 1) / TK * TK;
        splits = (K + ks - 1) / ks;
        if (splits < 1) splits = 1;
    }
    dim3 grid((M + TM - 1) / TM, (N + TN - 1) / TN, splits);
    float* part = splits > 1 ? d_work : nullptr;
    if (K == 0) part = nullptr;
    const float* init = (flags & ICESUR_ROWINIT) ? d_work : nullptr;
#define GEMM_LAUNCH(TA_, TB_) hipLaunchKernelGGL((k_gemm<TA_, TB_>), grid, dim3(256), 0, s, M, N, K, d_a, lda, d_b, ldb, E, ks > 0 ? ks : 1, part, init)
    if (!trans_a && !trans_b) GEMM_LAUNCH(false, false);
    else if (trans_a && !trans_b) GEMM_LAUNCH(true, false);
    else if (!trans_a && trans_b) GEMM_LAUNCH(false, true);
    else GEMM_LAUNCH(true, true);
#undef GEMM_LAUNCH
    if (part && splits > 32) {
        const int ng = (splits + 31) / 32;
        float* part2 = d_work + (size_t)splits * M * N;
        hipLaunchKernelGGL(k_splitk_group, dim3(grid1((size_t)M * N), ng), dim3(256), 0, s, M, N, splits, d_work, part2);
        hipLaunchKernelGGL(k_splitk_reduce, dim3(grid1((size_t)M * N)), dim3(256), 0, s, M, N, ng, part2, E);
    } else if (part) {
        hipLaunchKernelGGL(k_splitk_reduce, dim3(grid1((size_t)M * N)), dim3(256), 0, s, M, N, splits, d_work, E);
    }
    SUR_LAUNCHED();
    return ICEDEM_OK;
}

int icesur_conv(const uint16_t* d_x, int32_t B, int32_t H, int32_t W, int32_t C, const uint16_t* d_w, int32_t N,
                const float* d_bias, int32_t flags, uint16_t* d_wt, uint16_t* d_out, void* stream) {
    SUR_CHECK(B >= 0 && H > 0 && W > 0 && C > 0 && N > 0 && d_x && d_w && d_wt && d_out, "icesur_conv: bad arguments");
    SUR_CHECK(!(flags & ICESUR_BIAS) || d_bias, "icesur_conv: bias missing");
    SUR_CHECK(!(flags & ICESUR_ACCUM), "icesur_conv: accumulate is not supported");
    const int Ho = (H + 1) / 2, Wo = (W + 1) / 2, K = 9 * C;
    const size_t M = (size_t)B * Ho * Wo;
    SUR_CHECK(M < (size_t)1 << 31, "icesur_conv: too many output pixels");
    if (M == 0) return ICEDEM_OK;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_transpose16, dim3(grid1((size_t)K * N)), dim3(256), 0, s, d_w, K, N, d_wt);
    const int mode = ((C % 32 == 0) && ((uintptr_t)d_x % 16 == 0) && ((uintptr_t)d_wt % 16 == 0)) ? CONV_VEC
                     : (C == 2 && ((uintptr_t)d_x % 4 == 0)) ? CONV_PAIR : CONV_GEN;
    const int tn = (N % 128 == 0) ? 128 : (N % 64 == 0) ? 64 : 32;
    dim3 grid((unsigned)((M + CM - 1) / CM), (unsigned)((N + tn - 1) / tn));
#define CONV_LAUNCH(TN_, M_) hipLaunchKernelGGL((k_conv<TN_, M_>), grid, dim3(256), 0, s, d_x, B, H, W, C, Ho, Wo, d_wt, N, d_bias, d_out, flags)
#define CONV_MODES(TN_) do { if (mode == CONV_VEC) CONV_LAUNCH(TN_, CONV_VEC); else if (mode == CONV_PAIR) CONV_LAUNCH(TN_, CONV_PAIR); else CONV_LAUNCH(TN_, CONV_GEN); } while (0)
    if (tn == 128) CONV_MODES(128);
    else if (tn == 64) CONV_MODES(64);
    else CONV_MODES(32);
#undef CONV_MODES
#undef CONV_LAUNCH
    SUR_LAUNCHED();
    return ICEDEM_OK;
}

int icesur_conv12(const uint16_t* d_x, int32_t B, const uint16_t* d_w1, const float* d_b1, const uint16_t* d_w2,
                  const float* d_b2, int32_t flags, uint16_t* d_wt1, uint16_t* d_wt2, uint16_t* d_out, void* stream) {
    SUR_CHECK(B >= 0 && d_x && d_w1 && d_w2 && d_wt1 && d_wt2 && d_out, "icesur_conv12: bad arguments");
    SUR_CHECK(!(flags & ICESUR_BIAS) || (d_b1 && d_b2), "icesur_conv12: bias missing");
    SUR_CHECK(!(flags & ~(ICESUR_BIAS | ICESUR_RELU)), "icesur_conv12: only bias and relu");
    SUR_CHECK(((uintptr_t)d_x % 16) == 0 && ((uintptr_t)d_wt2 % 16) == 0 && ((uintptr_t)d_out % 16) == 0,
              "icesur_conv12: d_x, d_wt2 and d_out must be 16-byte aligned");
    SUR_CHECK(B < 65536, "icesur_conv12: B >= 65536");
    if (B == 0) return ICEDEM_OK;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_transpose16, dim3(grid1(18 * 32)), dim3(256), 0, s, d_w1, 18, 32, d_wt1);
    hipLaunchKernelGGL(k_transpose16, dim3(grid1(288 * 64)), dim3(256), 0, s, d_w2, 288, 64, d_wt2);
    constexpr int R2 = 4;   // conv-2 rows per workgroup: 30 KB of LDS, 5 workgroups per CU
    hipLaunchKernelGGL((k_conv12<R2>), dim3((unsigned)B, 16 / R2), dim3(256), 0, s, d_x, d_wt1, d_b1, d_wt2, d_b2,
                       flags, d_out);
    SUR_LAUNCHED();
    return ICEDEM_OK;
}

int icesur_im2col(const uint16_t* d_x, int32_t B, int32_t H, int32_t W, int32_t C, uint16_t* d_col, void* stream) {
    SUR_CHECK(B >= 0 && H > 0 && W > 0 && C > 0 && d_x && d_col, "icesur_im2col: bad arguments");
    const int Ho = (H + 1) / 2, Wo = (W + 1) / 2;
    const size_t total = (size_t)B * Ho * Wo * 9 * C;
    if (total == 0) return ICEDEM_OK;
    hipLaunchKernelGGL(k_im2col, dim3(grid1(total)), dim3(256), 0, (hipStream_t)stream, d_x, B, H, W, C, Ho, Wo, d_col);
    SUR_LAUNCHED();
    return ICEDEM_OK;
}

int icesur_col2im(const float* d_dcol, int32_t B, int32_t H, int32_t W, int32_t C, float* d_dx, void* stream) {
    SUR_CHECK(B >= 0 && H > 0 && W > 0 && C > 0 && d_dcol && d_dx, "icesur_col2im: bad arguments");
    const int Ho = (H + 1) / 2, Wo = (W + 1) / 2;
    const size_t total = (size_t)B * H * W * C;
    if (total == 0) return ICEDEM_OK;
    hipLaunchKernelGGL(k_col2im, dim3(grid1(total)), dim3(256), 0, (hipStream_t)stream, d_dcol, B, H, W, C, Ho, Wo, d_dx);
    SUR_LAUNCHED();
    return ICEDEM_OK;
}

int icesur_relu_back(float* d_g, int32_t ldg, const uint16_t* d_y, int32_t ldy, int32_t rows, int32_t cols,
                     uint16_t* d_gb, int32_t ldgb, void* stream) {
    SUR_CHECK(rows >= 0 && cols >= 0 && d_g && d_y, "icesur_relu_back: bad arguments");
    if ((size_t)rows * cols == 0) return ICEDEM_OK;
    hipLaunchKernelGGL(k_relu_back, dim3(grid1((size_t)rows * cols)), dim3(256), 0, (hipStream_t)stream, d_g, ldg, d_y,
                       ldy, rows, cols, d_gb, ldgb);
    SUR_LAUNCHED();
    return ICEDEM_OK;
}

int icesur_colsum(const float* d_x, int32_t ldx, int32_t rows, int32_t cols, float* d_out, int32_t accumulate,
                  float* d_work, void* stream) {
    SUR_CHECK(rows >= 0 && cols > 0 && d_x && d_out && d_work, "icesur_colsum: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    const int nblk = (rows + 255) / 256;
    if (nblk > 0)
        hipLaunchKernelGGL(k_colsum_part, dim3(nblk, (cols + 63) / 64), dim3(256), 0, s, d_x, ldx, rows, cols, d_work);
    hipLaunchKernelGGL(k_colsum_final, dim3((cols + 15) / 16), dim3(256), 0, s, d_work, nblk, cols, d_out, accumulate);
    SUR_LAUNCHED();
    return ICEDEM_OK;
}

int icesur_to_bf16(const float* d_x, int64_t n, uint16_t* d_out, void* stream) {
    SUR_CHECK(n >= 0 && d_x && d_out, "icesur_to_bf16: bad arguments");
    if (n == 0) return ICEDEM_OK;
    hipLaunchKernelGGL(k_to_bf16, dim3(grid1((size_t)n)), dim3(256), 0, (hipStream_t)stream, d_x, n, d_out);
    SUR_LAUNCHED();
    return ICEDEM_OK;
}

int icesur_adam(float* d_w, const float* d_g, float* d_m, float* d_v, int64_t n, float lr, float beta1, float beta2,
                float eps, int32_t t, uint16_t* d_wb, void* stream) {
    SUR_CHECK(n >= 0 && t >= 1 && d_w && d_g && d_m && d_v, "icesur_adam: bad arguments");
    if (n == 0) return ICEDEM_OK;
    const float c1 = 1.0f - powf(beta1, (float)t), c2 = 1.0f - powf(beta2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3(grid1((size_t)n)), dim3(256), 0, (hipStream_t)stream, d_w, d_g, d_m, d_v, n, lr,
                       beta1, beta2, eps, c1, c2, 0.0f, d_wb);
    SUR_LAUNCHED();
    return ICEDEM_OK;
}

int icesur_adamw(float* d_w, const float* d_g, float* d_m, float* d_v, int64_t n, float lr, float beta1,
                 float beta2, float eps, float weight_decay, int32_t t, uint16_t* d_wb, void* stream) {
    SUR_CHECK(n >= 0 && t >= 1 && weight_decay >= 0.0f && d_w && d_g && d_m && d_v, "icesur_adamw: bad arguments");
    if (n == 0) return ICEDEM_OK;
    const float c1 = 1.0f - powf(beta1, (float)t), c2 = 1.0f - powf(beta2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3(grid1((size_t)n)), dim3(256), 0, (hipStream_t)stream, d_w, d_g, d_m, d_v, n, lr,
                       beta1, beta2, eps, c1, c2, weight_decay, d_wb);
    SUR_LAUNCHED();
    return ICEDEM_OK;
}

int icesur_mse(const float* d_y, const float* d_t, const float* d_w, int32_t B, int32_t D, float* d_dy,
               uint16_t* d_dyb, float* d_loss, void* stream) {
    SUR_CHECK(B > 0 && D > 0 && d_y && d_t && d_w && d_dy && d_loss, "icesur_mse: bad arguments");
    hipLaunchKernelGGL(k_mse, dim3(1), dim3(256), 0, (hipStream_t)stream, d_y, d_t, d_w, B, D, d_dy, d_dyb, d_loss);
    SUR_LAUNCHED();
    return ICEDEM_OK;
}

int icesur_concat(const uint16_t* d_feat, int32_t ldf, int32_t F, const float* d_extra, int32_t lde, int32_t E,
                  int32_t B, uint16_t* d_z, int32_t ldz, void* stream) {
    SUR_CHECK(B >= 0 && F >= 0 && E >= 0 && ldz >= F + E && d_feat && d_z && (E == 0 || d_extra),
              "icesur_concat: bad arguments");
    if ((size_t)B * ldz == 0) return ICEDEM_OK;
    hipLaunchKernelGGL(k_concat, dim3(grid1((size_t)B * ldz)), dim3(256), 0, (hipStream_t)stream, d_feat, ldf, F,
                       d_extra, lde, E, B, d_z, ldz);
    SUR_LAUNCHED();
    return ICEDEM_OK;
}

int icesur_split(const float* d_dz, int32_t lddz, int32_t F, int32_t B, float* d_dfeat, void* stream) {
    SUR_CHECK(B >= 0 && F >= 0 && lddz >= F && d_dz && d_dfeat, "icesur_split: bad arguments");
    if ((size_t)B * F == 0) return ICEDEM_OK;
    hipLaunchKernelGGL(k_split, dim3(grid1((size_t)B * F)), dim3(256), 0, (hipStream_t)stream, d_dz, lddz, F, B, d_dfeat);
    SUR_LAUNCHED();
    return ICEDEM_OK;
}

int icesur_raster(const float* d_pos, const float* d_area, const float* d_thick, int32_t n, const float* d_pose,
                  int32_t B, int32_t G, float cell, float x_lo, int32_t* d_acc, uint16_t* d_out, void* stream) {
    SUR_CHECK(n >= 0 && B >= 0 && G > 0 && G <= 4096 && cell > 0.0f && d_pose && d_acc && d_out &&
              (n == 0 || (d_pos && d_area && d_thick)), "icesur_raster: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    if ((size_t)B * n > 0)
        hipLaunchKernelGGL(k_raster_splat, dim3(grid1((size_t)B * n)), dim3(256), 0, s, d_pos, d_area, d_thick, n,
                           d_pose, B, G, cell, x_lo, d_acc);
    if ((size_t)B * G * G > 0)
        hipLaunchKernelGGL(k_raster_out, dim3(grid1((size_t)B * G * G)), dim3(256), 0, s, d_acc, B, G, cell, d_out);
    SUR_LAUNCHED();
    return ICEDEM_OK;
}

int icesur_ke_costmap(const float* d_pos, const float* d_radius, const float* d_mass, int32_t n, int32_t rows,
                      int32_t cols, float x0, float y0, float res, double collision_cost, double ship_speed,
                      double ship_mass, double ke_cost_scale, float buffer_factor, float base_cost,
                      int32_t dilate_half, int32_t* d_work, float* d_out, void* stream) {
    SUR_CHECK(n >= 0 && rows > 0 && cols > 0 && res > 0.0f && dilate_half >= 0 && d_work && d_out &&
              (n == 0 || (d_pos && d_radius && d_mass)), "icesur_ke_costmap: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    const int cells = rows * cols;
    hipLaunchKernelGGL(k_ke_zero, dim3(grid1(cells)), dim3(256), 0, s, d_work, cells);
    if (n > 0)
        hipLaunchKernelGGL(k_ke_splat, dim3(grid1(n)), dim3(256), 0, s, d_pos, d_radius, d_mass, n, rows, cols, x0,
                           y0, res, collision_cost, fmax(ship_speed, 1e-6), ship_mass, ke_cost_scale, buffer_factor,
                           d_work);
    hipLaunchKernelGGL(k_ke_dilate, dim3(grid1(cells)), dim3(256), 0, s, d_work, rows, cols, base_cost, dilate_half,
                       d_out);
    SUR_LAUNCHED();
    return ICEDEM_OK;
}

int icesur_mbd_update(const float* d_costs, const float* d_samples, int32_t S, int32_t D, float temperature,
                      const float* d_ybar, float alpha_bar, float alpha, float alpha_bar_prev, float* d_work,
                      float* d_out, void* stream) {
    SUR_CHECK(S > 0 && D > 0 && temperature > 0.0f && d_costs && d_samples && d_ybar && d_work && d_out,
              "icesur_mbd_update: bad arguments");
    SUR_CHECK(alpha_bar > 0.0f && alpha_bar < 1.0f && alpha > 0.0f && alpha_bar_prev > 0.0f,
              "icesur_mbd_update: schedule values out of range");
    hipStream_t s = (hipStream_t)stream;
    const int nb = (S + 255) / 256;
    float* st = d_work;
    float* part = d_work + 2;
    hipLaunchKernelGGL(k_mppi_stats, dim3(1), dim3(256), 0, s, d_costs, S, temperature, st);
    hipLaunchKernelGGL(k_mppi_part, dim3(nb), dim3(256), 0, s, d_costs, d_samples, S, D, temperature, st, part);
    hipLaunchKernelGGL(k_mbd_final, dim3((D + 255) / 256), dim3(256), 0, s, part, nb, D, st, d_ybar, alpha_bar,
                       alpha, alpha_bar_prev, d_out);
    SUR_LAUNCHED();
    return ICEDEM_OK;
}

int icesur_mppi_update(const float* d_costs, const float* d_samples, int32_t S, int32_t D, float temperature,
                       float* d_work, float* d_out, void* stream) {
    SUR_CHECK(S > 0 && D > 0 && temperature > 0.0f && d_costs && d_samples && d_work && d_out,
              "icesur_mppi_update: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    const int nb = (S + 255) / 256;
    float* st = d_work;
    float* part = d_work + 2;
    hipLaunchKernelGGL(k_mppi_stats, dim3(1), dim3(256), 0, s, d_costs, S, temperature, st);
    hipLaunchKernelGGL(k_mppi_part, dim3(nb), dim3(256), 0, s, d_costs, d_samples, S, D, temperature, st, part);
    hipLaunchKernelGGL(k_mppi_final, dim3((D + 255) / 256), dim3(256), 0, s, part, nb, D, st, d_out);
    SUR_LAUNCHED();
    return ICEDEM_OK;
}

int icesur_raster_poly(const float* d_verts, const int32_t* d_nv, const float* d_thick, int32_t n, int32_t max_v,
                       const float* d_pose, int32_t B, int32_t G, float cell, float x_lo, int32_t* d_acc,
                       uint16_t* d_out, void* stream) {
    SUR_CHECK(n >= 0 && max_v >= 3 && B >= 0 && G > 0 && G <= 4096 && cell > 0.0f && d_pose && d_acc && d_out &&
              (n == 0 || (d_verts && d_nv && d_thick)), "icesur_raster_poly: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    if ((size_t)B * n > 0)
        hipLaunchKernelGGL(k_raster_poly, dim3(grid1((size_t)B * n)), dim3(256), 0, s, (const float2*)d_verts, d_nv,
                           d_thick, n, max_v, d_pose, B, G, cell, x_lo, d_acc);
    if ((size_t)B * G * G > 0)
        hipLaunchKernelGGL(k_raster_out, dim3(grid1((size_t)B * G * G)), dim3(256), 0, s, d_acc, B, G, cell, d_out);
    SUR_LAUNCHED();
    return ICEDEM_OK;
}

}  // extern "C"

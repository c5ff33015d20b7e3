/*
 * icesur.h -- C ABI of the rollout-cost surrogate kernels (gfx950, BF16 MFMA).
 *
 * No reference counterpart: the reference has no learned model of the DEM.
 * The surrogate (a CNN over a 64x64x2 ice raster, concatenated with the ship
 * state and the control sequence, then an MLP) is assembled by the Python
 * host (mbd-icenav_amd/icedem/surrogate.py) from these operations; every
 * floating-point operation of its forward and backward pass runs here.
 *
 * Conventions: bf16 tensors are uint16_t bit patterns (round to nearest
 * even), matrices are row-major with a leading dimension in elements, device
 * pointers d_*, `stream` a hipStream_t.  Every function returns 0 or an
 * ICEDEM_E_* code (icedem.h); icedem_last_error(NULL) describes it.
 * Reductions (split-K, column sums, col2im) run in a fixed order, so results
 * are identical from run to run.
 */
#ifndef ICESUR_H
#define ICESUR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ICESUR_RELU 1      /* epilogue: max(x, 0) */
#define ICESUR_ACCUM 2     /* epilogue: C += result (fp32 C only) */
#define ICESUR_BIAS 4      /* epilogue: + bias[col] */
#define ICESUR_ROWINIT 8   /* accumulators start from d_work[col] (N floats, every row) instead of 0; split_k 1 */

/* C (M x N) = op(A) (M x K) . op(B) (K x N) with bf16 operands and fp32
 * accumulation on MFMA; op(A) = A (M x K, lda) or, with trans_a, A^T where
 * A is K x M; op(B) = B (K x N, ldb) or, with trans_b, B^T where B is N x K.
 * The epilogue adds bias, applies ReLU and accumulates per `flags`, then
 * writes fp32 d_c (ldc) and/or bf16 d_cb (ldcb); either may be NULL.
 * split_k > 1 partitions K over split_k workgroup sets whose fp32 partial
 * products go to d_work and are added in split order before the epilogue;
 * with more than 32 splits, in split order within groups of 32 and then the
 * group sums in order.  d_work holds split_k * M * N floats, plus
 * ceil(split_k / 32) * M * N when split_k > 32.  With ICESUR_ROWINIT
 * (split_k 1) d_work instead holds one fp32 row of N accumulator start
 * values: a product whose first K0 columns of A are one row shared by all
 * rows equals, bit for bit, that row's product over those K0 (a 1-row GEMM,
 * K0 a multiple of 32) used as the start of the remaining K - K0. */
int icesur_gemm(int32_t trans_a, int32_t trans_b, int32_t M, int32_t N, int32_t K,
                const uint16_t* d_a, int32_t lda, const uint16_t* d_b, int32_t ldb,
                float* d_c, int32_t ldc, uint16_t* d_cb, int32_t ldcb, const float* d_bias, int32_t flags,
                int32_t split_k, float* d_work, void* stream);

/* 3x3 stride-2 padding-1 convolution of NHWC bf16 d_x (B, H, W, C) with
 * weights d_w (9 C, N) bf16 (row k = tap C + c, tap = 3 kh + kw), as an
 * implicit GEMM (no im2col matrix): d_out (B, Ho, Wo, N) bf16 =
 * [relu](conv + bias), Ho = (H + 1) / 2.  flags: ICESUR_BIAS, ICESUR_RELU.
 * d_wt: caller workspace of 9 C N bf16 (the transposed weights).  Equals
 * icesur_im2col followed by icesur_gemm with the same flags. */
int icesur_conv(const uint16_t* d_x, int32_t B, int32_t H, int32_t W, int32_t C, const uint16_t* d_w, int32_t N,
                const float* d_bias, int32_t flags, uint16_t* d_wt, uint16_t* d_out, void* stream);

/* The network's first two convolutions in one launch, for inference: d_x
 * (B, 64, 64, 2) bf16 -> conv (d_w1 (18, 32), d_b1) -> conv (d_w2 (288, 64),
 * d_b2) -> d_out (B, 16, 16, 64) bf16, each [relu](conv + bias) per flags
 * (ICESUR_BIAS, ICESUR_RELU).  Bit-identical to two icesur_conv calls; the
 * intermediate (B, 32, 32, 32) activation is not written.  d_wt1 (18 x 32)
 * and d_wt2 (288 x 64, 16-byte aligned) are caller workspace for the
 * transposed weights; d_x and d_out 16-byte aligned; B < 65536. */
int icesur_conv12(const uint16_t* d_x, int32_t B, const uint16_t* d_w1, const float* d_b1, const uint16_t* d_w2,
                  const float* d_b2, int32_t flags, uint16_t* d_wt1, uint16_t* d_wt2, uint16_t* d_out, void* stream);

/* 3x3, stride 2, padding 1 patches of an NHWC bf16 tensor (B, H, W, C):
 * d_col row b*Ho*Wo + ho*Wo + wo holds (kh*3 + kw)*C + c, Ho = (H+1)/2,
 * Wo = (W+1)/2, zeros outside the input. */
int icesur_im2col(const uint16_t* d_x, int32_t B, int32_t H, int32_t W, int32_t C, uint16_t* d_col,
                  void* stream);

/* Adjoint of icesur_im2col: d_dx (B, H, W, C) fp32 = the sum over the
 * patches that read each input position (kh, kw order). */
int icesur_col2im(const float* d_dcol, int32_t B, int32_t H, int32_t W, int32_t C, float* d_dx, void* stream);

/* ReLU backward: g[i] = d_y[i] > 0 ? g[i] : 0 in place (fp32; d_y is the
 * layer's bf16 output), and d_gb = bf16(g) when not NULL.  Rows of `cols`
 * values; g has row stride ldg, d_y ldy, d_gb ldgb. */
int icesur_relu_back(float* d_g, int32_t ldg, const uint16_t* d_y, int32_t ldy, int32_t rows, int32_t cols,
                     uint16_t* d_gb, int32_t ldgb, void* stream);

/* d_out[c] (+)= sum over rows of d_x[r * ldx + c] in a fixed order: rows in
 * blocks of 256, each block as 4 interleaved running sums (rows p, p + 4, ...)
 * added as (s0 + s1) + (s2 + s3); the block sums as 16 interleaved running
 * sums added as a pairwise tree.  d_work holds ceil(rows/256) * cols floats.
 * accumulate != 0 adds to d_out. */
int icesur_colsum(const float* d_x, int32_t ldx, int32_t rows, int32_t cols, float* d_out, int32_t accumulate,
                  float* d_work, void* stream);

/* bf16 copy of n floats (round to nearest even). */
int icesur_to_bf16(const float* d_x, int64_t n, uint16_t* d_out, void* stream);

/* Adam step t (>= 1) on n fp32 parameters d_w with gradients d_g and moments
 * d_m, d_v; d_wb (optional) receives bf16(w). */
int icesur_adam(float* d_w, const float* d_g, float* d_m, float* d_v, int64_t n, float lr, float beta1,
                float beta2, float eps, int32_t t, uint16_t* d_wb, void* stream);

/* Adam with decoupled weight decay: w -= lr (m_hat / (sqrt(v_hat) + eps) +
 * weight_decay w); otherwise as icesur_adam. */
int icesur_adamw(float* d_w, const float* d_g, float* d_m, float* d_v, int64_t n, float lr, float beta1,
                 float beta2, float eps, float weight_decay, int32_t t, uint16_t* d_wb, void* stream);

/* Weighted mean squared error over B x D predictions (ld = D): d_loss[0] =
 * sum_{b,o} w_o (y - t)^2 / (B D) (rows summed in order), d_dy = its gradient
 * 2 w_o (y - t) / (B D) in fp32 and, if d_dyb, bf16. */
int icesur_mse(const float* d_y, const float* d_t, const float* d_w, int32_t B, int32_t D, float* d_dy,
               uint16_t* d_dyb, float* d_loss, void* stream);

/* MLP input rows: d_z[b * ldz + j] = bf16 of feat[b, j] for j < F (the
 * flattened CNN output, bf16, ldf), then the E extra features of sample b
 * (fp32, lde), then zeros up to ldz. */
int icesur_concat(const uint16_t* d_feat, int32_t ldf, int32_t F, const float* d_extra, int32_t lde, int32_t E,
                  int32_t B, uint16_t* d_z, int32_t ldz, void* stream);

/* Gradient of the CNN part from the MLP input gradient: d_dfeat[b * F + j] =
 * d_dz[b * lddz + j] for j < F (fp32). */
int icesur_split(const float* d_dz, int32_t lddz, int32_t F, int32_t B, float* d_dfeat, void* stream);

/* Ice raster of a scene seen from a ship pose: a G x G grid of `cell` metres
 * in the ship frame (x forward, y to port), covering x in [x_lo, x_lo + G
 * cell) and y in [-G cell / 2, G cell / 2).  For each of the B poses
 * (d_pose[b] = x, y, heading), element e of the scene (centroid d_pos (2,n)
 * SoA, area d_area, thickness d_thick) is taken as the disc of its area
 * around its centroid and adds the area of every cell whose centre the disc
 * covers (fixed point, 2^-12 m^2); d_out (B, G, G, 2) bf16 gets the area
 * fraction (capped at 1) and the area-weighted mean thickness.  d_acc
 * (B G G 2 ints) must be zero on entry and is left zero. */
int icesur_raster(const float* d_pos, const float* d_area, const float* d_thick, int32_t n, const float* d_pose,
                  int32_t B, int32_t G, float cell, float x_lo, int32_t* d_acc, uint16_t* d_out, void* stream);

/* The KE collision costmap of the reference's planners (build_costmap_jax,
 * jax_costmap.py:177-330, ice_resistance_weight = 0): d_out (rows, cols),
 * row = y, origin (x0, y0), cell res, = max(base_cost, max over elements of
 * ke_i max(0, (r_e^2 - d^2) / r_e^2)) with r_e = r_i (1 + buffer_factor) +
 * res / 2 and ke_i = collision_cost v^2 m_i M (m_i + 2 M) / (2 (M + m_i)^2)
 * ke_cost_scale clipped to [0, 1e8] (v = ship_speed, M = ship_mass), then the
 * maximum over a (2 dilate_half + 1)^2 window with edge padding.  d_pos
 * (2, n) centroids, d_work rows * cols ints. */
int icesur_ke_costmap(const float* d_pos, const float* d_radius, const float* d_mass, int32_t n, int32_t rows,
                      int32_t cols, float x0, float y0, float res, double collision_cost, double ship_speed,
                      double ship_mass, double ke_cost_scale, float buffer_factor, float base_cost,
                      int32_t dilate_half, int32_t* d_work, float* d_out, void* stream);

/* One MPPI update of the reference's planners (run_mppi,
 * jax_mppi_planner.py:34-115): with beta = min(costs) and
 * w_i = exp(-(costs[i] - beta) / temperature), d_out (D) =
 * clip(sum_i w_i samples[i] / sum_i w_i, -1, 1) over d_samples (S, D).
 * Fixed-order sums; d_work holds 2 + ceil(S / 256) * D floats. */
int icesur_mppi_update(const float* d_costs, const float* d_samples, int32_t S, int32_t D, float temperature,
                       float* d_work, float* d_out, void* stream);

/* One reverse-diffusion step of the reference's MBD planner (run_mbd,
 * jax_mbd_planner.py:1363-1396): d_samples (S, D) are the step's clipped
 * samples Y0s, d_costs (S) their costs, d_ybar (D) the current mean.  With
 * run_mppi's weights w_i = exp(-(c_i - min c) / temperature),
 * Ybar_new = sum_i w_i Y0s_i / sum_i w_i, Yi = Ybar sqrt(alpha_bar),
 * score = (-Yi + sqrt(alpha_bar) Ybar_new) / (1 - alpha_bar),
 * d_out = clip((Yi + (1 - alpha_bar) score) / sqrt(alpha) / sqrt(alpha_bar_prev), -1, 1)
 * where alpha_bar = alphas_bar[i], alpha = alphas[i], alpha_bar_prev =
 * alphas_bar[max(i - 1, 0)].  Fixed-order sums; d_work holds
 * 2 + ceil(S / 256) * D floats; d_out may alias d_ybar. */
int icesur_mbd_update(const float* d_costs, const float* d_samples, int32_t S, int32_t D, float temperature,
                      const float* d_ybar, float alpha_bar, float alpha, float alpha_bar_prev, float* d_work,
                      float* d_out, void* stream);

/* The same window from the elements' convex polygons: d_verts (n, max_v, 2)
 * world vertices (the first d_nv[e] used, either orientation), each cell
 * sampled at 4 x 4 points; coverage and thickness in the same fixed point
 * as icesur_raster.  No reference counterpart. */
int icesur_raster_poly(const float* d_verts, const int32_t* d_nv, const float* d_thick, int32_t n, int32_t max_v,
                       const float* d_pose, int32_t B, int32_t G, float cell, float x_lo, int32_t* d_acc,
                       uint16_t* d_out, void* stream);

#ifdef __cplusplus
}
#endif
#endif

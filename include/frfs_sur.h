/*
 * frfs_sur.h -- C ABI of the learned stand-in for Q(f,f): a Fourier-layer
 * neural operator with its own forward, backward, loss and Adam kernels,
 * plus the AoSoA gather and the conservation check / scatter around it.
 *
 * The reference has no learned operator; the stand-in takes the place of
 * the same plugin call as the spectral operator,
 *   DGFSVHSGLLScatteringModel.fs (frfs/solvers/dgfs/scattering.py:227-286)
 *   DGFSBiVSSGLLScatteringModel.fs, one species
 *       (frfs/solvers/dgfsbi/scattering.py:323-330)
 * batched like fq_eval, and reads and writes the same AoSoA matrices
 * (frfs/backends/base/types.py:26-47; dof = elem*nupts + upt).  DOFs whose
 * conservation defect exceeds the tolerance are flagged by fq_sur_finish and
 * left for the spectral operator (fq_eval) to fill.
 *
 * Network (N^3 grid with N = 16 or 32, width 32, 8^3 retained modes,
 * 1..8 layers):
 *   v_0 = P g + p0;  z_l = Wm_l v_l + b_l + Re IDFT(R_l . DFT(v_l)/n);
 *   v_{l+1} = gelu(z_l) (last layer: z_l);  out = Qw . v_L + q0,
 * with g = f / (rho f_ref) and out ~ Q / (rho^2 q_ref) (Q is quadratic in f).
 * theta layout (floats): P (W), p0 (W), per layer [R (W*W*m^3 complex,
 * (i, o, k) order, interleaved), Wm (W*W, (o, i)), b (W)], Qw (W), q0.
 * Activation arrays are [batch][channel][x] FP32, x = (I*N + J)*N + K.
 *
 * All buffers are caller-owned device memory; nothing here allocates.
 * Errors: FQ_SUCCESS or a negative FQ_ERR_* code with fq_last_error().
 */
#ifndef FRFS_SUR_H
#define FRFS_SUR_H

#include <stdint.h>

#include "frfs_fq.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FQ_SUR_ABI_VERSION 1
#define FQ_SUR_MAX_LAYERS 8
#define FQ_SUR_MAX_BATCH 65536
#define FQ_SUR_FUSED_FLOATS 65697   /* folded inference weights (fq_sur_fuse) */
#define FQ_SUR_ACTS_FP32 0          /* activation storage of fq_sur_forward_fused */
#define FQ_SUR_ACTS_FP16 1
#define FQ_SUR_ACTS_BF16 2

typedef struct fq_sur_shape {
    int32_t abi_version;   /* FQ_SUR_ABI_VERSION */
    int32_t Nv;            /* grid points per axis: 16 or 32 */
    int32_t width;         /* channels: 32 */
    int32_t modes;         /* retained modes per axis: 8 */
    int32_t layers;        /* Fourier layers: 1..FQ_SUR_MAX_LAYERS */
} fq_sur_shape;

/* Sizes in floats: the parameter vector, the activations of a forward call
 * (training != 0 keeps what fq_sur_backward needs) and the backward work
 * array, for a batch of `batch` DOFs.  Any output pointer may be NULL. */
int fq_sur_sizes(const fq_sur_shape *shape, int64_t batch, int32_t training, int64_t *nparams,
                 int64_t *act_size, int64_t *work_size);

/* out[batch][N^3] = network(g[batch][N^3]).  acts: fq_sur_sizes floats.
 * acts_type (FQ_SUR_ACTS_*) is the storage of the activations; arithmetic
 * and the gradients stay FP32 (BF16 storage with FP32 accumulation for
 * training, for example).  A training call and the fq_sur_backward call that
 * uses its activations take the same acts_type. */
int fq_sur_forward(const fq_sur_shape *shape, const float *theta, const float *g, int64_t batch,
                   int32_t training, int32_t acts_type, float *acts, float *out, void *stream);

/* Inference with the lift folded into the first layer and the projection
 * into the last (layers >= 2): fq_sur_fuse computes the folded weights
 * (FQ_SUR_FUSED_FLOATS floats) from theta; fq_sur_forward_fused then gives
 * the output of fq_sur_forward up to rounding, transforming one channel
 * instead of 32 in the first and last layers.  acts: the inference size
 * (training = 0) of fq_sur_sizes.  Recompute the folded weights whenever
 * theta changes.  acts_type FQ_SUR_ACTS_FP16 or _BF16 stores the activations
 * between kernels in 16 bits (arithmetic stays FP32; half the activation
 * bytes; about 2e-3 (FP16) or 1e-2 (BF16) relative change of the output). */
int fq_sur_fuse(const fq_sur_shape *shape, const float *theta, float *fused, void *stream);
int fq_sur_forward_fused(const fq_sur_shape *shape, const float *theta, const float *fused, const float *g,
                         int64_t batch, int32_t acts_type, float *acts, float *out, void *stream);

/* grad += d(sum gout*out)/d theta, from the activations of a training
 * forward call on the same theta and g (grad is caller-zeroed; acts is
 * overwritten only in its scratch part). */
int fq_sur_backward(const fq_sur_shape *shape, const float *theta, const float *g, int64_t batch,
                    int32_t acts_type, float *acts, const float *gout, float *grad, float *work, void *stream);

/* gout = 2 (out - target) / count;  sums[0] += sum (out - target)^2,
 * sums[1] += sum target^2 (sums caller-zeroed, FP64). */
int fq_sur_loss(int64_t count, const float *out, const float *target, float *gout, double *sums,
                void *stream);

/* Conservation penalty added to the loss of fq_sur_loss (training only; no
 * reference counterpart): for each of the batch rows of out (the network
 * output on the Nv^3 grid on [-L, L]^3, cell volume cw), u = its orthogonal
 * projection onto span{1, cx, cy, cz, |c|^2} (gram_inv = the inverse of
 * that basis' Gram matrix in the cw inner product, 5 x 5 row major);
 * gout += 2 lambda u / count and sums[2] += |u|^2 (count = batch Nv^3), the
 * gradient and value of lambda |u|^2 / count.  sums holds three doubles. */
int fq_sur_cons(int32_t Nv, double L, double cw, double lambda, int64_t batch, const float *out,
                const double *gram_inv, float *gout, double *sums, void *stream);

/* One Adam step (bias-corrected, step >= 1) over the whole parameter vector. */
int fq_sur_adam(int64_t nparams, float *theta, const float *grad, float *m, float *v, float lr,
                float beta1, float beta2, float eps, int32_t step, void *stream);

/* dst[b][x] = src(first_dof + b, x) / (rho_b^power * scale) from an AoSoA
 * matrix; rho_b = rho_in[b], or (rho_in NULL) the density cw * sum_x src,
 * written to rho_out when that is not NULL. */
int fq_sur_gather(int32_t Nv, double cw, double scale, int32_t power, const double *src,
                  int32_t nupts, int64_t ldim, int32_t neles, int64_t first_dof, int64_t ndof,
                  const double *rho_in, double *rho_out, float *dst, void *stream);

/* Conservation check and scatter of a normalised network output q[ndof][n]:
 * defect_b = max_j |<q, phi_j>| / (|phi_j| |q|) over the collision invariants
 * phi = (1, cx, cy, cz, |c|^2) in the cw inner product (gram_diag = |phi_j|^2,
 * gram_inv the inverse Gram matrix, row-major 5x5).  flag_b = defect_b > tol;
 * unflagged DOFs get Q = rho_b^2 scale (q - projection of q onto the
 * invariants) written into dst; flagged DOFs are left untouched. */
int fq_sur_finish(int32_t Nv, double L, double cw, double scale, double tol, const float *q,
                  const double *rho, const double *gram_inv, const double *gram_diag, double *dst,
                  int32_t nupts, int64_t ldim, int32_t neles, int64_t first_dof, int64_t ndof,
                  double *defect, int32_t *flag, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FRFS_SUR_H */

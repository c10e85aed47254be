/*
 * samrai_learned.h -- C ABI of the learned block smoother in libsamrai_fac.so.
 *
 * None of these entry points has a counterpart in the reference: SAMRAI's
 * CellPoissonFACOps smooths only with red-black Gauss-Seidel
 * (solv/CellPoissonFACOps.C:1338-1514, called through
 * FACOperatorStrategy::smoothError, solv/FACOperatorStrategy.h:82-364).
 * The learned smoother is an extra correction a smoothError implementation
 * may apply per 16^3 block of a level before (or instead of) red-black
 * sweeps; the red-black smoother remains the fallback and the FP64 residual
 * remains the convergence test.
 *
 * Blocks: 16x16x16 cells, x fastest, blocks contiguous (block b, cell
 * (x,y,z) at b*4096 + (z*16 + y)*16 + x).  Face coefficients of a block:
 * dxf[b][z][y][0..16], dyf[b][z][0..16][x], dzf[b][0..16][y][x].
 * BF16 tensors are raw uint16 bit patterns, channels-last with 16 channels
 * per cell.  All pointers are device pointers; every call is
 * stream-ordered, returns 0 on success and < 0 on error, with the message
 * in sfac_last_error() (samrai_fac.h).
 */
#ifndef SAMRAI_LEARNED_H
#define SAMRAI_LEARNED_H

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#include "samrai_fac.h"

#ifdef __cplusplus
extern "C" {
#endif

/* 3x3x3 convolution, 16 -> 16 channels, zero padding at block faces, BF16
 * operands on the matrix cores with FP32 accumulation.  w[t][n][k] with
 * t = 9(dz+1) + 3(dy+1) + (dx+1), n the output and k the input channel.
 * mode 0: out = relu(conv(x) + bias) (BF16);
 * mode 1: pred[cell] = conv(x)[cell][0] + bias[0] (FP32);
 * mode 2: out = conv(x) * (mask > 0) (BF16; backward of a ReLU layer when w
 *         is the tap-flipped transposed weight tensor). */
int sfac_conv3x3x3_bf16(const uint16_t* x, const uint16_t* w, const float* bias,
                        const uint16_t* mask, uint16_t* out, float* pred, int nblocks, int mode,
                        hipStream_t stream);

/* Weight gradient of one convolution layer:
 * grad[t][co][ci] = sum_cells gz[cell][co] * x[cell + off(t)][ci] (t < 27),
 * grad[27][co][0] = sum_cells gz[cell][co] (bias), other entries of tap 27
 * are sums against a zero column.  partial: device scratch of
 * sfac_conv3x3x3_wgrad_partial_size(nblocks) floats. */
int sfac_conv3x3x3_wgrad_bf16(const uint16_t* x, const uint16_t* gz, float* grad, float* partial,
                              int nblocks, hipStream_t stream);
size_t sfac_conv3x3x3_wgrad_partial_size(int nblocks);

/* No counterpart in the reference: the weight gradient above for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * or training loops, harnesses that forbid hipMalloc).  It allocates nothing:
 * the per-wave slabs live in a static device array of the library, one per
 * GPU, 7 MiB, held until the process ends; its address is looked up for the
 * current device on the first call there (the stream must belong to the
 * current device).  nblocks <= SFAC_WGRAD_NOALLOC_MAX_BLOCKS, larger counts
 * are rejected with an error.  Results are bit-identical to
 * sfac_conv3x3x3_wgrad_bf16.  Every call on one GPU shares that array, so
 * calls that overlap in time on one GPU (several streams or host threads)
 * race; calls ordered on one stream do not. */
#define SFAC_WGRAD_NOALLOC_MAX_BLOCKS 16
int sfac_conv3x3x3_wgrad_bf16_noalloc(const uint16_t* x, const uint16_t* gz, float* grad,
                                      int nblocks, hipStream_t stream);

/* Block problems div(D grad e) = r, homogeneous Dirichlet on the block faces
 * (ghost = -interior), FP64.  sfac_block_sor: red-black SOR sweeps (omega 1
 * is red-black Gauss-Seidel), starting from e (from_input != 0) or zero.
 * sfac_block_residual: rout = r - div(D grad e). */
int sfac_block_sor(const double* r, const double* dxf, const double* dyf, const double* dzf,
                   double h, int sweeps, double omega, int from_input, double* e, int nblocks,
                   hipStream_t stream);
int sfac_block_residual(const double* r, const double* dxf, const double* dyf, const double* dzf,
                        double h, const double* e, double* rout, int nblocks, hipStream_t stream);

/* The network: 2 -> 16 -> 16 -> 1 channels, 3x3x3 convolutions, ReLU after
 * the hidden layers.  Input features per cell: r h^2 / s and log D, with s
 * the RMS of r h^2 over the block; output e / s.  FP32 master parameters
 * and Adam state on the device, BF16 copies for the matrix cores.
 * Parameters of layer l (3 layers, 27*16*16 + 16 floats each): w[t][co][ci]
 * then bias[co]; gradients per layer: 28*16*16 floats as in
 * sfac_conv3x3x3_wgrad_bf16. */
int sfac_learned_create(int max_blocks, unsigned long long seed, void** net);
int sfac_learned_destroy(void* net);
int sfac_learned_features(void* net, const double* r, const double* logd, double h, int nblocks,
                          hipStream_t stream);
int sfac_learned_forward(void* net, int nblocks, hipStream_t stream);
/* e = omega * s * prediction (FP64), after features + forward. */
int sfac_learned_correction(void* net, double omega, double* e, int nblocks, hipStream_t stream);
/* Loss mean((prediction - target / s)^2) and the gradients of every layer,
 * after features + forward on the same blocks. */
int sfac_learned_backward(void* net, const double* target, int nblocks, hipStream_t stream);
int sfac_learned_adam(void* net, float lr, hipStream_t stream);
/* features + forward + backward + Adam. */
int sfac_learned_train_step(void* net, const double* r, const double* logd, const double* target,
                            double h, int nblocks, float lr, hipStream_t stream);
/* which 0: parameters (3 * (27*16*16 + 16) floats); 1: gradients (3 * 28*16*16). Synchronous. */
int sfac_learned_get_params(void* net, float* host, int which);
int sfac_learned_set_params(void* net, const float* host);
int sfac_learned_loss(void* net, float* loss);
/* Device buffers of the last features/forward: features (BF16, 16 channels per
 * cell), prediction (FP32 per cell), per-block scale s (FP64). */
int sfac_learned_buffers(void* net, void** features, float** pred, double** scale);

/* The learned correction on level ln of a solver's hierarchy (cubic
 * cells): each patch is covered by 16x16x16 windows from its lower corner
 * (one network block each; window cells past the patch read 0 and are not
 * corrected), and u += omega * s * net(r h^2 / s, log D) on every patch
 * interior.  r: cell data with ghost 0 (for example from sfac_ops_residual);
 * D: the side data given to the solver (ghost 0), or NULL for log D = 0; u:
 * cell data of any ghost width (ghosts are not touched: fill them before the
 * next smoothing).  Runs on the solver's stream; the network must hold at
 * least as many blocks as the level has windows (a 16x16x16 patch is one).
 * A smoothError implementation calls it before red-black sweeps. */
int sfac_ops_learned_correct(sfac_solver* s, void* net, sfac_patch_data* u, sfac_patch_data* r,
                             sfac_patch_data* D, int ln, double omega);

/* Attach the network to a solver (net NULL detaches): CellPoissonFACOps::
 * smoothError then applies omega times the learned correction of the error
 * equation's residual before its red-black sweeps, on every level above the
 * coarsest with cubic cells (patches of any size, covered by 16x16x16
 * windows as above).  Each solveSystem starts with it enabled and falls back
 * to red-black alone for the rest of the solve after a cycle whose residual
 * reduction factor exceeds 0.5; the FP64 residual remains the convergence
 * test.  The network must hold at least as many blocks as such a level has
 * windows (checked when attached to an initialized solver). */
int sfac_solver_set_learned_smoother(sfac_solver* s, void* net, double omega);
/* Which smoothing calls of the recursive V-cycle apply it: bit 0 the
 * pre-smoothing, bit 1 the post-smoothing (default 3, both). */
int sfac_solver_set_learned_phases(sfac_solver* s, int mask);
/* 1 if the learned correction is attached and enabled (it is switched off
 * by the fallback), 0 if not, -1 for a null solver. */
int sfac_solver_learned_active(sfac_solver* s);
/* Cycles of the last solveSystem that the fallback undid (0 or 1: after
 * one the correction stays off for the rest of that solve), -1 for a null
 * solver. */
int sfac_solver_learned_rollbacks(sfac_solver* s);

#ifdef __cplusplus
}
#endif

#endif

/*
 * dgpu_closure.h — C ABI of the learned LES closure: a small MLP
 * (13 -> 64 -> 64 -> 2, GELU) from invariants of the resolved velocity and
 * buoyancy gradients plus the filter width to the eddy viscosity and
 * diffusivity, trained and evaluated by hand-written BF16 MFMA kernels with FP32
 * accumulation.
 *
 * It has no counterpart in the reference.  It is a further LES model beside the
 * reference's les_model_type 1, 4, 5 (les_chan / les_chan_th, for/les.f90:3-1435):
 * les_model_type 6 evaluates the network at the GY points, where type 5 (AMD)
 * evaluates its eddy viscosity and diffusivity, takes the mean of two GY
 * neighbours at the GYF points between them, and falls back to the type-5
 * coefficient wherever a point's features leave the training envelope (at a GYF
 * point: where either neighbour does).  Conventions as in diablo_gpu.h (int status, synchronous
 * calls, the library owns device memory, dgpu_last_error for the message).
 *
 * Network:  y = W3^T g(W2^T g(W1^T x' + b1) + b2) + b3,  g = GELU in its tanh form
 *   g(x) = x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))/2,
 * x' = (x - shift) * scale per feature.  Parameter block (float, DGPU_MLP_NPARAM):
 *   W1 [16][64] (rows 13..15 zero), b1 [64], W2 [64][64], b2 [64],
 *   W3 [64][16] (columns 2..15 zero), b3 [16] (entries 2..15 zero)
 * with W stored [in][out].  Normaliser (float, DGPU_MLP_NNORM):
 *   shift [16], scale [16], lo [16], hi [16]   (lo/hi: the training envelope of
 *   the raw features; padded entries 13..15: shift 0, scale 0, lo -inf, hi +inf).
 * Features (raw, float [16] per point, entries 13..15 zero), from the
 * scaled gradient tensor G' of the type-5 closure (G'_ij = du_i/dx_j delta_j /
 * delta_i, for/les.f90:571-780) with strain S and rotation O, the scaled
 * buoyancy gradient b (delta_k dth/dx_k), s2 = S:S, w2 = O:O, e = s2 + w2:
 *   0 (w2 - s2)/e                 1 tr(S^3)/s2^1.5           2 tr(S O O^T)/(s2^.5 w2)
 *   3 -tr(G' G'^T S)/(s2^.5 e)    4 tr(S^2 O O^T)/(s2 w2)     5 b_y^2/|b|^2
 *   6 b.S.b/(|b|^2 s2^.5)          7 b.S^2.b/(|b|^2 s2)        8 b.O O^T.b/(|b|^2 w2)
 *   9 tanh(g_y th_y / s2)         10 tanh(|g_y| |grad th| / s2)   11 log(1 + D2 sqrt(2 s2)/nu)
 *  12 log(delta_y^2 / (delta_x delta_z))
 * where D2 = 3/(1/delta_x^2 + 1/delta_y^2 + 1/delta_z^2) (the type-5 filter
 * width, for/les.f90:756-780), th the first scalar and g_y = grav_y, its factor in
 * the wall-normal momentum equation (for/solver.f90:518-554).  A term whose
 * denominator vanishes is 0 (no scalar, |b| = 0, w2 = 0).  Outputs: y0, y1 with
 *   nu_t = C^2 D2 sqrt(2 s2) max(0, y0),  kappa_t = C^2 D2 sqrt(2 s2) max(0, y1),
 * C = 0.2887 (the type-5 constant, for/les.f90:25).
 */
#ifndef DGPU_CLOSURE_H
#define DGPU_CLOSURE_H

#include "diablo_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
  DGPU_MLP_NFEAT = 13,
  DGPU_MLP_NFEAT_PAD = 16,
  DGPU_MLP_HIDDEN = 64,
  DGPU_MLP_NOUT = 2,
  DGPU_MLP_NPARAM = 16 * 64 + 64 + 64 * 64 + 64 + 64 * 16 + 16,
  DGPU_MLP_NNORM = 64
};

typedef struct dgpu_mlp dgpu_mlp;

/* A network on the device.  params: DGPU_MLP_NPARAM floats, or NULL for a seeded
 * initialisation (Glorot-uniform weights, zero biases, the output layer scaled
 * by 0.1).  The normaliser starts as shift 0, scale 1, envelope unbounded. */
int dgpu_mlp_create(const float *params, uint64_t seed, dgpu_mlp **out);
int dgpu_mlp_destroy(dgpu_mlp *m);

/* Parameters in/out (FP32 master copy; the kernels read BF16 copies of the
 * weights).  Setting the parameters resets the Adam moments and step count. */
int dgpu_mlp_set_params(dgpu_mlp *m, const float *params);
int dgpu_mlp_get_params(dgpu_mlp *m, float *params);
int dgpu_mlp_set_norm(dgpu_mlp *m, const float *norm);

/* Training data, copied to the device: X [n][16] raw features, T [n][2]
 * targets, 1 <= n < 2^31. */
int dgpu_mlp_set_data(dgpu_mlp *m, const float *X, const float *T, int64_t n);

/* Full-batch gradient of L = (1/n) sum_s sum_o w_o (y_so - t_so)^2 over the
 * first n samples of the data: grad (DGPU_MLP_NPARAM floats) and *loss. */
int dgpu_mlp_grad(dgpu_mlp *m, int64_t n, const float loss_w[2], float *grad, float *loss);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with the given gradient. */
int dgpu_mlp_adam_step(dgpu_mlp *m, const float *grad, float lr);

/* Minibatch Adam over the data for `epochs` epochs (a fresh permutation per
 * epoch from `seed`; the last minibatch of an epoch may be short).
 * epoch_loss (epochs floats, may be NULL): the mean loss of each epoch, taken
 * at the parameters each minibatch saw. */
int dgpu_mlp_train(dgpu_mlp *m, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                   const float loss_w[2], float *epoch_loss);

/* Inference on raw features X [n][16] -> Y [n][2] (before the max(0, .)). */
int dgpu_mlp_predict(dgpu_mlp *m, const float *X, int64_t n, float *Y);

/* Give an LES handle (les_model_type 6) its network: parameters and normaliser
 * are copied; later changes to a dgpu_mlp do not reach the handle. */
int dgpu_set_closure(dgpu_handle *h, const float *params, const float *norm);

/* Of the GY points of rows 2..Nyp the closure evaluated in the last LES
 * substep (*points = (Ny - 1) Nz Nx), those whose velocity coefficient came from
 * the network (*network); the rest took the AMD fallback.  Meaningful after an
 * LES substep of the handle.  No counterpart in the reference (a statistic of
 * the learned closure). */
int dgpu_closure_counts(dgpu_handle *h, int64_t *network, int64_t *points);

/* Features of the current state at the GYF points of rows jstart..Nyp-1 (the
 * rows where les_chan forms the GYF stresses), after the BC posts of an LES
 * substep at `time`: X [n][16], scale [n] = C^2 D2 sqrt(2 s2), amd [n][2] = the
 * type-5 velocity and scalar coefficients at the point divided by sqrt(2 s2) (so
 * nu_t(AMD) = scale * amd[.][0], kappa_t(AMD) = scale * amd[.][1]),
 * inside [n] = 1 where every feature lies in the handle's envelope (needs a
 * closure set; else all 1).  n = (Nyp - jstart) * Nz * Nx points, ordered
 * [j][z][x].  *n_out receives n; any output pointer but n_out may be NULL. */
int dgpu_les_features(dgpu_handle *h, double time, float *X, float *scale, float *amd, int32_t *inside,
                      int64_t *n_out);

/* Training data: the handle's state (a DNS) box-filtered onto a coarser grid of
 * the same domain, with the products of the subgrid stress and flux.  No
 * counterpart in the reference's solver (its post-processing filters DNS
 * offline, proc/functions/coarsen.m:1-25).  In x and z a top-hat of width fx dx
 * (fz dz) centred on every fx-th (fz-th) point, trapezoid weights over its
 * fx + 1 (fz + 1) points, periodic (fx, fz even divisors of Nx, Nz); in y the
 * rows of an LES grid with nl interior rows: mf [nl+2][Ny+2] weights of the DNS
 * rows for the GYF fields, mg [nl+2][Ny+2] for the GY field (row-major; the
 * caller's choice, e.g. the overlap of the LES cells with the DNS cells).  out:
 * [14][nl+2][Nz/fz][Nx/fx] doubles, the filtered
 *   0 u1, 1 u2 (mg), 2 u2 at GYF, 3 u3, 4 th,
 *   5 u1u1, 6 u1u2, 7 u1u3, 8 u2u2, 9 u2u3, 10 u3u3, 11 u1 th, 12 u2 th, 13 u3 th
 * with u2 at GYF the DNS's own (u2(j) + u2(j+1))/2 on rows 1..Ny, th scalar 1.
 * Needs n_th >= 1.  The state is not changed. */
int dgpu_box_filter(dgpu_handle *h, int32_t fx, int32_t fz, int32_t nl, const double *mf, const double *mg,
                    double *out);

#ifdef __cplusplus
}
#endif
#endif /* DGPU_CLOSURE_H */

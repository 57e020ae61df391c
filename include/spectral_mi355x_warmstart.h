/*
 * spectral_mi355x_warmstart.h -- the learned warm start.
 *
 * No counterpart in the reference: its LOBPCG start block is N(0,1) draws plus
 * the sqrt-degree hint (src/solvers/lobpcg.rs:243-252, also the seeds of its
 * warm restarts, :380-390).  A 3-layer message-passing network (hidden 64,
 * tanh; node features: degree, local normalised-weight statistics, 8
 * random-walk probes) maps every node to 8 outputs, trained so that their span
 * covers the two non-trivial eigenvectors.  The warm-start block is the
 * Rayleigh-Ritz block of [D^{1/2} 1 | the outputs for three independent probe
 * sets]; it reaches a solver through si_config.warm_start, and only when its
 * residual ||L x - lambda x|| beats a random start's.
 * Dense layers run on the bf16 matrix cores, training uses the project's own
 * forward and backward kernels and Adam on the device.
 *
 * Blocks cross this ABI in the caller's node order, on the device.
 */
#ifndef SPECTRAL_MI355X_WARMSTART_H
#define SPECTRAL_MI355X_WARMSTART_H

#include "spectral_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct si_mpnn si_mpnn;

/* number of FP32 parameters (packed: per layer Ws, Wn, b; then Wo [64 x 8]) */
int64_t si_mpnn_param_count(void);
/* new network with Glorot-uniform weights drawn from seed */
si_status si_mpnn_create(si_context* ctx, uint64_t seed, si_mpnn** out);
void si_mpnn_destroy(si_mpnn* m);
/* host copies of the FP32 master parameters (count = si_mpnn_param_count()) */
si_status si_mpnn_get_params(si_context* ctx, const si_mpnn* m, float* out, int64_t count);
si_status si_mpnn_set_params(si_context* ctx, si_mpnn* m, const float* in, int64_t count);
/* node features f [n x 16] (device, FP32) of a Laplacian; probes drawn from seed */
si_status si_mpnn_features(si_context* ctx, const si_laplacian* lap, uint64_t seed, float* f);
/* network output x [n x 8] (device, FP32) for features f */
si_status si_mpnn_forward(si_context* ctx, si_mpnn* m, const si_laplacian* lap, const float* f,
                          float* x);
/* coverage loss tr(V^T V) - tr(B^T (G + 1e-3 tr(G)/8 I)^-1 B), G = X^T X,
 * B = X^T V, against the target eigenvectors v [n x 2] (device, FP64), and its
 * gradient (host, FP64, packed like the parameters; may be NULL) */
si_status si_mpnn_gradient(si_context* ctx, si_mpnn* m, const si_laplacian* lap, const float* f,
                           const double* v, double* loss, double* grad);
/* one Adam step (lr) on the loss against v; loss before the step */
si_status si_mpnn_train_step(si_context* ctx, si_mpnn* m, const si_laplacian* lap, const float* f,
                             const double* v, float lr, double* loss);
/* warm-start block [n x 3] (device, FP64) for n_components = 2: the three
 * lowest Rayleigh-Ritz vectors of [D^{1/2} 1 | X(seed) | X(seed+1) | X(seed+2)]
 * (X(s) = the outputs for probes drawn from s) or, if their residual is not
 * smaller, the Rayleigh-Ritz block of the solver's random start from seed;
 * used_net / res_net / res_random (may be NULL) report the choice */
si_status si_mpnn_warm_start(si_context* ctx, si_mpnn* m, const si_laplacian* lap, uint64_t seed,
                             int32_t n_components, double* block, int32_t* used_net,
                             double* res_net, double* res_random);

#ifdef __cplusplus
}
#endif
#endif

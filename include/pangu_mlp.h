/*
 * pangu_mlp: learned starting bracket for the Kastaun primitive recovery,
 * and learned Newton start for the 1Dvsq recovery (no reference
 * counterpart).
 *
 * scheme_kastaun.h:110-165 searches the root mu of its master function by
 * false position on the fixed bracket [0, 1], with no initial guess.  This
 * module trains a small network, 5 -> 64 -> 64 -> 64 -> 1 with SiLU, to
 * predict mu from five dimensionless invariants of the conserved state, and
 * lets the recovery kernel start from a one-sided bracket [g, g (1 +- delta)]
 * on the side of the prediction g where the master function changes sign,
 * widened when it holds no root and replaced by [0, 1] as a last resort.  Weights
 * and activations enter the matrix products as BF16 on gfx950 MFMA units
 * with FP32 accumulation; FP32 master weights are trained with Adam.
 *
 * Features of a state (from D, q = tau/D, r^2 = S^2/D^2, b^2 = B^2/D and
 * (r.b)^2 of the Kastaun setup): log q, log r^2, log b^2, log (r.b)^2 (each
 * floored at 1e-12, 1e-24, 1e-24, 1e-48) and (r.b)^2 / (r^2 b^2 + 1e-300)
 * clamped to [0, 1], standardised with the training set's mean and
 * deviation.
 * Target: logit(mu); the prediction is g = 1 / (1 + exp(-z)).  A model of
 * kind 1 (pangu_mlp_set_kind) predicts z = log u^2 instead, and the 1Dvsq
 * iteration starts from v^2 = 1 / (1 + exp(-z)).
 */
#ifndef PANGU_MLP_H
#define PANGU_MLP_H
#include <stdint.h>

#include "pangu_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { PANGU_MLP_IN = 5, PANGU_MLP_HIDDEN = 64 };

typedef struct pangu_mlp pangu_mlp;

/* Parameter layout of pangu_mlp_get/set_params (FP32, this order):
 * W1[5][64], b1[64], W2[64][64], b2[64], W3[64][64], b3[64], W4[64], b4,
 * feature mean[5], feature deviation[5]: PANGU_MLP_NPARAM floats. */
enum { PANGU_MLP_NPARAM = 5 * 64 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1 + 5 + 5 };

int pangu_mlp_create(int device, uint64_t seed, pangu_mlp **out);
int pangu_mlp_destroy(pangu_mlp *m);
const char *pangu_mlp_last_error(const pangu_mlp *m);
int pangu_mlp_get_params(pangu_mlp *m, float *out);
int pangu_mlp_set_params(pangu_mlp *m, const float *in);

/* What the model predicts (before any data set is made): kind 0 (default)
 * the Kastaun root as logit(mu), for the learned bracket of the robuster
 * recovery; kind 1 log u^2 = log(W^2 - 1), the start of the 1Dvsq Newton
 * iteration (v^2 = u^2 / (1 + u^2)) of the onedvsq recovery, which
 * otherwise starts from the previous primitives' v^2
 * (scheme_1d_vsq.h:274-301).  Same features, network and trainer. */
int pangu_mlp_set_kind(pangu_mlp *m, int kind);

/* Synthetic recovery states on the GPU (the training / evaluation set):
 * rho ~ logU(1e-6, 1), u/rho ~ logU(1e-4, 10), W ~ U(1, 20),
 * sigma = b^2/rho ~ logU(1e-3, 1e2), random directions of u~ and B, Kerr
 * metric in MKS coordinates (a, h) at r ~ U(1.3, 50), x2 ~ U(-0.98, 0.98),
 * gamma as given.  Each state is converted to conserved variables and
 * solved by the plain Kastaun recovery; states it solves (status 0) are
 * kept with their root as target (kind 1: log(W^2 - 1) of the drawn state,
 * which both solvers recover).  Returns the number kept in *kept.  A new
 * call replaces the previous set. */
int pangu_mlp_make_dataset(pangu_mlp *m, int64_t n, uint64_t seed, double a, double h,
                           double gamma, int64_t *kept);

/* Trains on the current data set: `epochs` passes in minibatches of `batch`
 * (a multiple of 32) with Adam, the learning rate decaying from lr to
 * lr / 100 along a cosine over the epochs of the call; the feature
 * standardisation is fixed from the data set on the first call.
 * loss_out (optional, epochs floats) receives each epoch's mean loss. */
int pangu_mlp_train(pangu_mlp *m, int epochs, int batch, float lr, float *loss_out);

/* One forward and backward pass on host data (n a multiple of 32): feats
 * [n][5] raw features, target [n]; returns the loss and the FP32 gradient
 * of every parameter (PANGU_MLP_NPARAM floats; the statistics' entries are
 * zero).  For gradient checks. */
int pangu_mlp_gradient(pangu_mlp *m, const float *feats, const float *target, int n, float *loss,
                       float *grad);

/* Forward pass on host features [n][5]; out[n] receives g in (0, 1). */
int pangu_mlp_predict(pangu_mlp *m, const float *feats, int n, float *out);

/* Recovery of the current data set by plain Kastaun and by the learned
 * bracket (delta as in pangu_set_c2p_guess): mean false-position
 * iterations and root evaluations of each, the largest |f(mu)| of the
 * learned roots, the largest relative difference of the recovered rho, u
 * and u~ between the two, and the number of states the learned path failed
 * on.  out: {plain_iters, learned_iters, plain_evals, learned_evals,
 * max_residual, max_rel_diff, failures}.  Kind 1: 1Dvsq Newton iterations
 * from v^2 = 0 (no previous state) against the learned start, with entropy
 * U_ent / D; max_residual is the largest |f| / (Q~^2 + (B^2 + W)^2 v^2) of
 * the 1Dvsq residual (func_1d_gnr) at the learned results, and the
 * evaluation counts equal the iteration counts. */
int pangu_mlp_evaluate(pangu_mlp *m, double delta, double *out);

/* C = A B for host FP32 matrices A [M][K] and B [K][N] (row-major), rounded
 * to BF16 and multiplied on the matrix cores with FP32 accumulation (the
 * product every layer uses).  M, N multiples of 32, K of 16. */
int pangu_mlp_gemm(int device, int M, int N, int K, const float *A, const float *B, float *C);

/* Appends training pairs from a running simulation to the data set: every
 * stride-th interior cell of container md of ctx (its conserved state in
 * the recovery's normalisation and its metric), solved by the plain Kastaun
 * recovery (kind 1: then by 1Dvsq from Kastaun's v^2, with entropy
 * U_ent / D); solved states are kept with their root (kind 1: log u^2) as
 * target (*kept receives how many).  Waits for the device first.  The data
 * set's states must share the context's adiabatic index. */
int pangu_mlp_harvest(pangu_mlp *m, pangu_ctx *ctx, int md, int stride, int64_t *kept);

/* The in-kernel inference path on host features [n][5]: the network as the
 * recovery kernel evaluates it, one wavefront over 64 states with the
 * activations kept in the matrix-core fragments (pangu_mlp_predict runs the
 * training forward pass instead).  out[n] receives g in (0, 1). */
int pangu_mlp_predict_wave(pangu_mlp *m, const float *feats, int n, float *out);

/* Attaches the model to a context.  Kind 0 (robuster recovery): Kastaun
 * then starts from the learned one-sided bracket of initial relative width
 * delta.  Kind 1 (onedvsq recovery): the 1Dvsq Newton iteration starts from
 * the learned v^2 (delta unused but must lie in (0, 1)).  NULL detaches.
 * The model must outlive the attachment.  The recovery kernel evaluates
 * the network itself, in the wavefronts the gate selects
 * (pangu_set_c2p_gate).  Runs with a model attached are off the bit-for-bit
 * parity path (results agree to the solver tolerance). */
int pangu_set_c2p_guess(pangu_ctx *ctx, pangu_mlp *m, double delta);

/* Gate of the learned bracket: a wavefront (64 cells of the recovery
 * launch) evaluates the network when one of its cells took at least
 * min_prev_iters false-position steps in its previous recovery (the
 * iteration count in c2p_iters, which the context must have); every cell
 * of that wavefront then starts from the learned bracket.  Wavefronts with
 * no such cell skip the network and run the plain search.  0 (the default)
 * uses it for every cell whose count is at hand. */
int pangu_set_c2p_gate(pangu_ctx *ctx, int min_prev_iters);

#ifdef __cplusplus
}
#endif
#endif /* PANGU_MLP_H */

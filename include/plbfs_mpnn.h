/*
 * plbfs_mpnn.h -- C-ABI of the warm-start message-passing network on the cell
 * graph (no counterpart in the reference, which always starts from its initial
 * conditions, src/Solution.cpp:94-211).
 *
 * 6 layers, hidden width 64, every MLP Linear -> SiLU -> Linear; products on the
 * matrix cores in bfloat16 with FP32 accumulation, FP32 master weights and Adam.
 * Inputs are padded to 16 per node (x) and per directed edge (f); outputs are
 * 16 per node.  Parameters are one flat FP32 vector: for each MLP in the order
 * enc, msg0, upd0, ..., msg5, upd5, dec the arrays W1 [k_in x 64] (row-major),
 * b1 [64], W2 [64 x k_out], b2 [k_out]; k_in = 16, 144, 128 (enc, msg, upd)
 * and 64 (dec), k_out = 64 except dec (16).
 */
#ifndef PLBFS_MPNN_H
#define PLBFS_MPNN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct plbfs_mpnn plbfs_mpnn;

/* Number of FP32 parameters. */
int plbfs_mpnn_n_params(void);
/* Graph of n_nodes nodes and n_edges directed edges src -> dst with 16 edge
 * features each ([n_edges*16]); train != 0 keeps every layer's activations for
 * the backward pass. */
int plbfs_mpnn_create(int32_t n_nodes, int32_t n_edges, const int32_t *src, const int32_t *dst,
                      const float *edge_feat, int32_t train, plbfs_mpnn **out);
int plbfs_mpnn_destroy(plbfs_mpnn *net);
/* Sets the weights (and clears the Adam moments) / reads them back. */
int plbfs_mpnn_set_params(plbfs_mpnn *net, const float *params);
int plbfs_mpnn_get_params(plbfs_mpnn *net, float *params);
/* y [n_nodes*16] = network(x [n_nodes*16]). */
int plbfs_mpnn_forward(plbfs_mpnn *net, const float *x, float *y);
/* Mean squared error over the outputs where mask != 0, and its gradient with
 * respect to every parameter (grads [n_params], optional). */
int plbfs_mpnn_loss_grad(plbfs_mpnn *net, const float *x, const float *target, const float *mask, double *loss,
                         float *grads);
/* One Adam step on the gradient of the last plbfs_mpnn_loss_grad. */
int plbfs_mpnn_adam_step(plbfs_mpnn *net, float lr, float beta1, float beta2, float eps);
/* epochs passes over n_samples (x, target) pairs held on the device, one Adam
 * step per sample in a seeded shuffled order; the step size follows a cosine
 * from lr in the first epoch down to lr/20 in the last.  loss_hist[epochs] gets
 * the mean loss of each epoch.  The sample copies are freed on return. */
int plbfs_mpnn_train(plbfs_mpnn *net, int32_t n_samples, const float *X, const float *T, const float *mask,
                     int32_t epochs, float lr, uint64_t seed, double *loss_hist);

#ifdef __cplusplus
}
#endif

#endif /* PLBFS_MPNN_H */

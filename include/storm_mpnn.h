/*
 * C ABI of the warm-start network (libstorm_mpnn.so): an edge
 * message-passing network over the cells of a reach that predicts an initial
 * state (h, hu, hv) for the FVT solver from the bed, the roughness and the
 * boundary discharge and stage.  No reference counterpart: SToRM starts from
 * the user's initial conditions (dataprepFVT.f90:90-130); the prediction is
 * handed to storm_fvt_set_state, and the FP64 solver then runs to the same
 * residual as from a cold start.
 *
 * Network (hidden width H, L layers), every product a BF16 MFMA GEMM with
 * FP32 accumulation:
 *   h0      = relu(x W_enc + b_enc)                         node encoder
 *   m_e     = relu([h_src, h_dst, e] W_e^l + b_e^l)        message of edge e
 *   a_i     = sum of m_e over the edges into i (fixed order)
 *   h^{l+1} = h^l + relu([h^l, a] W_n^l + b_n^l)           node update
 *   y       = h^L W_dec + b_dec                             decoder
 * Loss: mean over nodes and outputs of w_i (y - t)^2.  Training: AdamW.
 * Every sum is formed in a fixed order: results are bit-repeatable.
 *
 * Initial weights: He-uniform (seeded), zero biases, and a zero decoder,
 * so an untrained network outputs zero.
 *
 * Parameters are FP32 master weights, laid out in this order: W_enc
 * [f_node][H], b_enc[H]; per layer W_e [2H + f_edge][H], b_e[H], W_n [2H][H],
 * b_n[H]; W_dec [H][f_out], b_dec[f_out] (row-major [in][out]).  f_node,
 * f_edge and f_out must be multiples of 16; H must be 64.
 */
#ifndef STORM_MPNN_H
#define STORM_MPNN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct storm_mpnn_cfg {
  int32_t layers, hidden, f_node, f_edge, f_out;
  float lr, beta1, beta2, eps, weight_decay;
  uint64_t seed;  /* initial weights (uniform, He scale) */
} storm_mpnn_cfg_t;

typedef struct storm_mpnn storm_mpnn_t;

int storm_mpnn_create(const storm_mpnn_cfg_t *cfg, storm_mpnn_t **net);
void storm_mpnn_destroy(storm_mpnn_t *net);
int64_t storm_mpnn_num_params(const storm_mpnn_t *net);
int storm_mpnn_get_params(storm_mpnn_t *net, float *p);
int storm_mpnn_set_params(storm_mpnn_t *net, const float *p);
/* The graph of one sample: edges src[e] -> dst[e] (both directions of every
 * cell adjacency are listed; at least one edge), node features [n][f_node],
 * edge features [e][f_edge], both row-major FP32. */
int storm_mpnn_set_graph(storm_mpnn_t *net, int32_t n_nodes, int32_t n_edges, const int32_t *src,
                         const int32_t *dst, const float *node_feat, const float *edge_feat);
/* Forward pass; out (may be NULL) receives y [n][f_out]. */
int storm_mpnn_forward(storm_mpnn_t *net, float *out);
/* Loss of the last forward pass against target [n][f_out] with node weights
 * w[n] (NULL = 1), and the gradient of every parameter (overwritten). */
int storm_mpnn_backward(storm_mpnn_t *net, const float *target, const float *w, double *loss);
/* Gradients of the last backward pass, in the parameter layout. */
int storm_mpnn_get_grads(storm_mpnn_t *net, float *g);
/* One AdamW update with the gradients of the last backward pass. */
int storm_mpnn_adamw_step(storm_mpnn_t *net);
/* Learning rate of the following AdamW updates (a schedule's current value). */
int storm_mpnn_set_lr(storm_mpnn_t *net, float lr);
const char *storm_mpnn_last_error(storm_mpnn_t *net);

/* The depth profile of the non-learned start the network refines (host
 * code, no device): the steady 1-D gradually varied flow with unit
 * discharge q over a bed of slope S,
 *   dh/ds = -(S - Sf) / (1 - Fr^2),  Fr^2 = q^2 / (g h^3),
 *   Sf = manning^2 q^2 / h^(10/3),  g = 9.81,
 * integrated upstream from h[0] = h0 by the classical fourth-order
 * Runge-Kutta method on n_steps steps of ds, each stage's depth and each
 * result held at h_floor or above.  h: n_steps + 1 values.  Every operation
 * is the IEEE double one in the order written, with the C library's pow.
 * Returns 0, or -1 for a null h or a negative n_steps. */
int storm_mpnn_gvf_profile(double q, double S, double manning, double h0, double h_floor,
                           double ds, int32_t n_steps, double *h);

#ifdef __cplusplus
}
#endif
#endif

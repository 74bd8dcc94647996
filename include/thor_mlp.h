/*
 * thor_mlp.h — per-column MLP for the radiative-heating surrogate.
 *
 * No reference counterpart: THOR computes the heating with its double-gray
 * module only (rtm_dual_band, src/physics/modules/inc/profx_RT.h:296-587).
 * This network stands in for that column solve: per column, n_in inputs ->
 * n_hidden GELU -> n_hidden GELU -> n_out outputs, BF16 weights and
 * activations, FP32 accumulation on the matrix cores (gfx950 MFMA), FP32
 * master parameters updated with Adam. Training and inference run on the GPU
 * with the library's own kernels.
 *
 * Parameter layout (FP32, one flat array): W1[n_in][n_hidden], b1[n_hidden],
 * W2[n_hidden][n_hidden], b2[n_hidden], W3[n_hidden][n_out], b3[n_out].
 */
#ifndef THOR_MLP_H
#define THOR_MLP_H

#include "thor_dyn.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct thor_mlp thor_mlp;

/* Creates the network on the current HIP device with seeded Kaiming-uniform
 * weights and zero biases. */
thor_status thor_mlp_create(int n_in, int n_hidden, int n_out, unsigned long long seed, thor_mlp **out);
void thor_mlp_destroy(thor_mlp *m);
long thor_mlp_param_count(const thor_mlp *m);
thor_status thor_mlp_get_params(thor_mlp *m, float *dst, long n);
thor_status thor_mlp_set_params(thor_mlp *m, const float *src, long n);

/* Training set in host memory, row-major and already normalised:
 * X[n][n_in], Y[n][n_out]. Copied to the device once. */
thor_status thor_mlp_set_data(thor_mlp *m, const float *X, const float *Y, long n);
/* One epoch of Adam (beta1 0.9, beta2 0.999) over seeded shuffled
 * mini-batches of the training set; *loss = mean squared error over it. */
thor_status thor_mlp_train_epoch(thor_mlp *m, int batch, float lr, unsigned long long seed, double *loss);

/* Forward pass on host arrays (X[n][n_in] -> Y[n][n_out]). */
thor_status thor_mlp_forward(thor_mlp *m, const float *X, float *Y, long n);
/* Mean-squared loss and its gradient with respect to every parameter (flat
 * layout above) for one batch in host memory (n <= 65536). */
thor_status thor_mlp_gradients(thor_mlp *m, const float *X, const float *Y, long n, float *grad, double *loss);

/* Milliseconds of device time for one forward pass over n device-resident
 * rows (timing of the inference kernels alone). */
thor_status thor_mlp_time_forward(thor_mlp *m, long n, int repeats, double *ms);

/* --------------------------------------------------------- RT surrogate */

/* Normalisation and envelope of the surrogate. Inputs per column: x = [T_k
 * (nv), ln p_k (nv), cos z]; x_norm = (x - in_mean) / in_std. Output:
 * Qheat_k = y_k * out_std_k + out_mean_k [W/m3]. A column whose normalised
 * inputs leave [in_lo, in_hi] (element-wise) is out of the envelope and gets
 * the exact double-gray RT. Every judge_every-th surrogate ProfX runs the
 * exact RT on every column instead and records the surrogate's error. */
typedef struct thor_rt_surrogate_params {
    const float *in_mean, *in_std;   /* [2 nv + 1] */
    const float *in_lo, *in_hi;      /* [2 nv + 1] */
    const float *out_mean, *out_std; /* [nv] */
    int judge_every;                 /* 0: never */
} thor_rt_surrogate_params;

enum { THOR_FORCING_RT_SURROGATE = 4 };

/* Attaches a trained network (n_in = 2 nv + 1, n_out = nv) to a dynamical
 * core that already has thor_dyn_set_rt parameters (the fallback and judge).
 * The network must outlive the handle's use of it. Stream order between the
 * two is kept by the library: the network's entry points that rewrite its
 * weights or buffers (set_params, set_data, train_epoch, forward, gradients,
 * time_forward) first wait for the work every attached handle has queued,
 * and a surrogate ProfX waits for the network's queued work. */
thor_status thor_dyn_set_rt_surrogate(thor_dyn *h, thor_mlp *m, const thor_rt_surrogate_params *p);
/* Surrogate statistics since the last call: ProfX calls, judge calls,
 * columns that fell back to the exact RT, and the RMS and max absolute
 * Qheat error [W/m3] of the surrogate against the exact RT over the judged
 * columns. Resets the counters. */
thor_status thor_dyn_surrogate_stats(thor_dyn *h, long *calls, long *judged, long *fallback_columns,
                                     double *rms_err, double *max_err);

#ifdef __cplusplus
}
#endif
#endif /* THOR_MLP_H */

/*
 * C ABI of the learned warm start for the ionosphere potential solve.
 *
 * No counterpart in the reference: the reference starts each polar solve
 * from the previous potential (NO/nonorthogonal_mi.c:449-464,
 * cubedsphere_mi_dipole.c:407-412). With a network attached, the device
 * solve starts instead from
 *     previous potential + N(residual of the previous potential),
 * where N is a 4-level U-Net (32 base channels, bf16 activations, fp32
 * accumulation on the matrix cores) trained to map a field-aligned current
 * pattern to its potential. The Krylov solve that follows is unchanged and
 * runs to the same tolerance, so the potential it returns is the solver's.
 *
 * The network may run on the polar grid pooled f x f (its grid is
 * colatitude_points / f x azimuth_points / f): the Krylov iterations remove
 * short-wavelength error quickly, the network supplies the long-wavelength
 * part, and on a coarse grid a 4-level U-Net's receptive field spans the
 * polar cap. Its answer is interpolated back bilinearly (rows clamped,
 * longitude periodic).
 *
 * Network input features (32 channels), per point of its grid: the f x f
 * fine-grid currents of the point's block / the current's RMS in channels
 * 0 .. f*f - 1 (row major, f <= 4), SigmaP / 10, SigmaH / 10, colatitude /
 * maximum colatitude and hemisphere sign (-1 north, +1 south) in channels
 * 16 .. 19, zeros elsewhere. Output: the block mean of the potential /
 * (R^2 RMS / SigmaP), divided by the network's output scale.
 *
 * Every function returns 0 on success and -1 on failure (gcs_last_error()).
 */
#ifndef GAMERA_CS_UNET_H
#define GAMERA_CS_UNET_H

#include <stddef.h>
#include <stdint.h>

#include "gamera_cs_step.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gcs_unet gcs_unet;

/* Training on generated pairs: each step draws `batch` random smooth
 * potentials with uniform conductances in the given ranges, computes the
 * current for which each solves the problem's polar system, and takes one
 * Adam step on the network's relative error (per sample, the mean square
 * error over the mean square target, averaged over the batch). The
 * learning rate decays by a cosine to 5% after a 50-step linear warmup. */
typedef struct {
  long steps;
  int batch;
  double learning_rate;
  unsigned long long seed;
  double pedersen_min, pedersen_max; /* S */
  double hall_min, hall_max;         /* S */
} gcs_warm_train_config;

/* A network for a colatitude_points x azimuth_points grid (both multiples
 * of 8) and up to max_batch samples, He-normal weights from `seed`. */
gcs_unet *gcs_unet_create(int colatitude_points, int azimuth_points, int max_batch,
                          unsigned long long seed);
void gcs_unet_destroy(gcs_unet *net);

/* Parameters, packed layer by layer: for each 3x3 convolution the weights
 * [cout][3][3][cin] then the bias [cout], then the 1x1 head (32 weights,
 * one bias). */
size_t gcs_unet_param_count(const gcs_unet *net);
int gcs_unet_set_params(gcs_unet *net, const float *params, size_t count);
int gcs_unet_get_params(gcs_unet *net, float *params, size_t count);
/* Gradients of the last gcs_unet_gradients call, same packing. */
int gcs_unet_get_grads(gcs_unet *net, float *grads, size_t count);
double gcs_unet_output_scale(const gcs_unet *net);
int gcs_unet_set_output_scale(gcs_unet *net, double scale);

/* Forward pass: features [batch][H][W][32] -> out [batch][H][W]. */
int gcs_unet_forward(gcs_unet *net, const float *features, int batch, float *out);
/* Forward, mean-square-error loss against target [batch][H][W], backward. */
int gcs_unet_gradients(gcs_unet *net, const float *features, const float *target,
                       int batch, double *loss);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) with the stored gradients. */
int gcs_unet_adam_step(gcs_unet *net, double learning_rate);

/* Check of the matrix-core operand layout: c[m][n] = a[m][k] b[k][n] with
 * bf16 a, b (raw 16-bit patterns) and fp32 c; m, n, k multiples of 16. */
int gcs_mfma_gemm_check(const uint16_t *a, const uint16_t *b, int m, int n, int k,
                        float *c);

/* Attach (or, with NULL, detach) the network to the problem's MI solve. The
 * network grid must be the MI grid pooled by one factor f in both
 * directions (f = 1: the MI grid itself) and its batch at least 2. */
int gcs_problem_mi_attach_warm_start(gcs_problem *problem, gcs_unet *net);
/* Train on pairs for the problem's MI grid; loss_history (steps entries, may
 * be NULL) receives each step's loss. */
int gcs_problem_mi_train_warm_start(gcs_problem *problem, gcs_unet *net,
                                    const gcs_warm_train_config *cfg,
                                    float *loss_history);
/* Train on given pairs: current and potential [count][nt][np] (each one
 * hemisphere, hemisphere[b] = GAMERA_MI_NORTH or GAMERA_MI_SOUTH), e.g.
 * currents logged from MHD runs (gcs_problem_mi_sample_current) with the
 * potentials the plain solve gives for them, or the differences of two
 * such pairs. Uses cfg's steps, batch, learning rate and seed, the same
 * loss and schedule as gcs_problem_mi_train_warm_start, and the problem's
 * conductances; the pairs stay on the device for the run. */
int gcs_problem_mi_train_warm_start_on(gcs_problem *problem, gcs_unet *net,
                                       const gcs_warm_train_config *cfg,
                                       const double *current, const double *potential,
                                       const int *hemisphere, int count,
                                       float *loss_history);
/* `count` generated pairs for the problem's MI grid with uniform
 * conductances: current and potential [count][nt][np]; sample b is in the
 * northern hemisphere for even b, the southern for odd b. */
int gcs_problem_mi_make_pairs(gcs_problem *problem, unsigned long long seed, int count,
                              double pedersen, double hall, double *current,
                              double *potential);
/* One polar solve of both hemispheres for current [2][nt][np] (north,
 * south), starting from potential [2][nt][np] and overwriting it with the
 * solution; with use_network the attached network corrects the guess
 * first. iterations[2] and milliseconds describe the solve. */
int gcs_problem_mi_solve_current(gcs_problem *problem, const double *current,
                                 double *potential, int use_network, int iterations[2],
                                 double *milliseconds);

#ifdef __cplusplus
}
#endif

#endif

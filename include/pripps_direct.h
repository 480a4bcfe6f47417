/*
 * pripps_direct.h -- C ABI of the MI355X (gfx950) Direct Fourier transform.
 *
 * The Direct scheme evaluates, per reciprocal-lattice q-vector of a periodic
 * box, the structure-factor amplitudes
 *
 *     A(q) = sum_i fA_i(|q|) exp(i q.r_i)   (atoms)
 *     E(q) = sum_i fE_i(|q|) exp(i q.r_i)   (excluded-volume dummies)
 *     H(q) = sum_i w_i fW(|q|) exp(i q.r_i) (hydration dummies)
 *
 * with q = 2 pi (n_x / L_x, n_y / L_y, n_z / L_z) for integer lattice indices
 * n, and optionally the per-molecule self term sum_m |A_m(q)|^2 over
 * contiguous equal-size molecules.  It replaces the reference's CPU
 * DirectTransform (src/explicit.rs:151-266: vacuum_rows, build_model) and its
 * wgpu kernel (src/gpu.rs:48-150, structure_factor), whose results are then
 * averaged over groups of equal |q| (average_by_qnorm, src/explicit.rs:399-420;
 * QGrid::group, src/q_grid.rs:133-157).
 *
 * The structure descriptor is pripps_debye_input (pripps_debye.h) with
 * nq = the number of q-vectors and q[v] = |q_v|: every form-factor table is
 * indexed by vector.  The q-vector lattice, box and grouping arrive in a
 * pripps_direct_grid.  Conventions (status codes, pripps_last_error, borrowed
 * inputs, caller-owned outputs, immutable models) are those of pripps_debye.h.
 */
#ifndef PRIPPS_DIRECT_H
#define PRIPPS_DIRECT_H

#include <stdint.h>

#include "pripps_debye.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Periodic box, q-vectors and their grouping into output rows. */
typedef struct {
    double box[3];            /* box side lengths L (Angstrom), > 0 */
    int32_t n_vectors;        /* q-vectors; must equal the input's nq */
    const int32_t *indices;   /* [n_vectors][3] lattice indices n; q = 2 pi n / L */
    int32_t num_molecules;    /* 0: whole box, no self term; M > 0: the atoms form
                                 contiguous molecules of max(n_atoms / M, 1) atoms (the
                                 last one shorter if M does not divide n_atoms) and the
                                 self term is computed; vacuum only (reference
                                 `with_num_molecules` and `chunks(m)`,
                                 src/explicit.rs:189-193, 217-244) */
    int32_t n_groups;         /* output rows */
    const int32_t *group;     /* [n_vectors] row of each vector, in [0, n_groups) */
    const double *group_q;    /* [n_groups] |q| reported for each row */
} pripps_direct_grid;

typedef struct pripps_direct_model pripps_direct_model;

/* Build a cached Direct model: amplitudes A, E, H (and the self term) of every
 * q-vector, resident on the device.  precision is PRIPPS_FP32_COMPENSATED
 * (FP32 phase and hardware sin/cos, FP32 sums over at most 256 atoms carried
 * into FP64) or PRIPPS_FP64 (FP64 throughout).
 * Replaces explicit::build_model (src/explicit.rs:283-326) and the vacuum rows
 * of DirectTransform (src/explicit.rs:217-244) / GpuDirectTransform
 * (src/gpu.rs:360-453). */
int pripps_direct_model_build(pripps_engine *engine, const pripps_debye_input *input,
                              const pripps_direct_grid *grid, int32_t precision,
                              pripps_direct_model **out);

/* Output rows and their |q| (reference DirectModel::q_values, src/explicit.rs:354-356). */
int pripps_direct_model_n_groups(const pripps_direct_model *model, int32_t *n_groups);
int pripps_direct_model_q_values(const pripps_direct_model *model, double *out);

/* Group means of |A - vs E + cd H|^2, |A|^2, |E|^2, |H|^2 (each [n_groups];
 * any output pointer may be NULL).  Reference DirectModel::intensity
 * (src/explicit.rs:328-351) with intensity_from_5col (src/explicit.rs:432-466).
 * For a vacuum model total = atoms = the reference's I_box. */
int pripps_direct_model_intensity(const pripps_direct_model *model, double volume_scale,
                                  double contrast_density, double *out_total,
                                  double *out_atoms, double *out_excluded,
                                  double *out_hydration);

/* Group means of the per-molecule self term sum_m |A_m|^2 ([n_groups]); the
 * model must have been built with num_molecules > 0.  Reference
 * Intensity::self_term of the vacuum rows (src/explicit.rs:207-214). */
int pripps_direct_model_self_term(const pripps_direct_model *model, double *out);

/* Group means of the six products |A|^2, |E|^2, |H|^2, Re A E*, Re A H*,
 * Re E H* ([6][n_groups], the pripps_debye.h profile order).  The intensity
 * is their Debye combine (DebyeTerms::combine, src/debye.rs:186-195), so
 * pripps_debye_model_from_profiles turns them into a model that
 * pripps_debye_model_fit can fit (reference DirectModel's fit, which closes
 * over DirectModel::intensity: pripps-py/src/lib.rs:625-650, 673-678). */
int pripps_direct_model_profiles(const pripps_direct_model *model, double *out);

/* Per-vector amplitudes [n_vectors][7]: Re A, Im A, Re E, Im E, Re H, Im H,
 * self term (0 without molecules). */
int pripps_direct_model_amplitudes(const pripps_direct_model *model, double *out);

/* Timing of the last build or recompute (kernel_ms, total_ms; pair_q_terms
 * holds atom-vector terms, n_atoms * n_vectors; work_items the launch's items). */
int pripps_direct_model_stats(const pripps_direct_model *model, pripps_build_stats *out);

/* Re-run the amplitude kernels `repeats` times on the resident inputs
 * (benchmarking); *kernel_ms receives the mean amplitude-kernel time. */
int pripps_direct_model_recompute(pripps_direct_model *model, int32_t repeats,
                                  double *kernel_ms);

int pripps_direct_model_free(pripps_direct_model *model);

/* Weighted average of I(q) over trajectory frames for the Direct scheme
 * (reference trajectory_average, src/trajectory.rs:167-251, with XtcIterator's
 * stride and weights, src/trajectory.rs:29-135).  `topology` gives kinds,
 * tables and weights (its x/y/z are not read); coords is
 * [n_frames][n_atoms][3] in Angstrom; frames 0, stride, 2 stride, ... are
 * used, with frame_weights (NULL: 1 each) indexed by the original frame.  The
 * box is fixed (the reference accumulates against the first frame's q grid).
 * out_total[n_groups] receives the weighted mean of group means of
 * |A - vs E + cd H|^2 (vs, cd baked in per frame, as the one-shot path does);
 * out_self[n_groups] (or NULL) the weighted mean self term, which needs
 * num_molecules > 0.  Positions are re-staged per frame into the resident
 * model; amplitudes and accumulators never leave the device. */
int pripps_direct_frames_average(pripps_engine *engine, const pripps_debye_input *topology,
                                 const pripps_direct_grid *grid, int32_t n_frames,
                                 const double *coords, const double *frame_weights,
                                 int32_t stride, int32_t precision, double volume_scale,
                                 double contrast_density, double *out_total, double *out_self);

#ifdef __cplusplus
}
#endif

#endif /* PRIPPS_DIRECT_H */

/*
 * wavefront_hip.h -- C ABI of the MI355X (gfx950) timestep engine.
 *
 * Drop-in boundary: the body of the reference's engine::run loop
 * (src/engine.rs:747-782), i.e. step_all_slabs (src/engine.rs:291-422) +
 * Source::inject (src/source.rs:79-95) per source + Probe::accumulate
 * (src/probe.rs:126-138) per probe, plus the snapshot serializer
 * (src/engine.rs:622-640).  The reference has no FFI of its own; these are the
 * entry points a Rust `-sys` crate binds.
 *
 * Conventions: extern "C" only; no exceptions cross the boundary; every
 * function returning int returns WF_OK (0) or a negative WF_ERR_* code and
 * leaves a message for wf_last_error() (thread-local).  Host buffers are
 * caller-owned; device buffers are owned by the wf_ctx.  A context is not
 * thread-safe; separate contexts may be used from separate threads.
 *
 * Field state on the host side always uses the reference snapshot format:
 * block-major 8x8x8 AoSoA, per block ex[512] ey[512] ez[512] hx[512] hy[512]
 * hz[512] little-endian f32, local index (lz*8+ly)*8+lx, blocks ordered
 * (bz*nby+by)*nbx+bx (src/layout.rs:697-787).  On the device the fields are
 * six x-fastest SoA arrays.
 */
#ifndef WAVEFRONT_HIP_H
#define WAVEFRONT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WF_OK 0
#define WF_ERR_INVALID (-1) /* bad argument or precondition */
#define WF_ERR_HIP (-2)     /* HIP runtime error (incl. no device) */
#define WF_ERR_OOM (-3)     /* device or host allocation failed */
#define WF_ERR_STATE (-4)   /* call out of order (e.g. step past source table) */
#define WF_ERR_IO (-5)      /* snapshot file I/O (incl. short writes) */

#define WF_BLOCK_DIM 8

/* Field components, src/probe.rs:30-38 (sources use 0..2 only). */
#define WF_EX 0
#define WF_EY 1
#define WF_EZ 2
#define WF_HX 3
#define WF_HY 4
#define WF_HZ 5

/* Waveform kinds, src/source.rs:30-43. */
#define WF_WAVE_GAUSSIAN 0 /* p0 = t0, p1 = spread */
#define WF_WAVE_SINUSOID 1 /* p0 = freq_hz */
#define WF_WAVE_RICKER 2   /* p0 = peak_freq_hz, p1 = t0 */

typedef struct wf_ctx wf_ctx;

/* A point source site.  Replaces Source's location/component
 * (src/source.rs:65-72).  hard != 0 assigns instead of adding (the sheet
 * source of examples/convergence_study.rs:165-173). */
typedef struct wf_source_site {
    uint32_t x, y, z;
    int32_t component; /* WF_EX..WF_EZ */
    int32_t hard;
} wf_source_site;

/* A full source description for the host-side evaluator
 * (src/source.rs:65-72 + src/main.rs:431-460 derive t0/spread). */
typedef struct wf_source_desc {
    wf_source_site site;
    float amplitude;
    int32_t waveform; /* WF_WAVE_* */
    float p0, p1;
} wf_source_desc;

/* A probe site (src/probe.rs:63-75 minus the accumulators). */
typedef struct wf_probe_site {
    uint32_t x, y, z;
    int32_t component; /* WF_EX..WF_HZ */
    uint32_t nfreq;
} wf_probe_site;

/* ---------------------------------------------------------------------- */
/* Errors / info                                                           */
/* ---------------------------------------------------------------------- */
const char *wf_last_error(void);
/* Number of visible HIP devices (0 without a GPU). */
int wf_device_count(void);

/* ---------------------------------------------------------------------- */
/* Host-side table builders (same arithmetic as the reference)              */
/* ---------------------------------------------------------------------- */
/* FixedQ16_16::from_f32 / to_f32, src/layout.rs:114-125 */
int32_t wf_q16_from_f32(float v);
float wf_q16_to_f32(int32_t q);
/* MaterialTable::coeffs_from_physical, src/layout.rs:194-208.
 * out = {ca, cb, da, db} in Q16.16. */
void wf_material_coeffs(float eps_r, float mu_r, float sigma, float dt, float d,
                        int32_t out[4]);
/* MaterialTable::vacuum_filled, src/layout.rs:175-180: table[256*4]. */
void wf_material_table_vacuum(float dt, float d, int32_t *table);
/* effective_pml_thickness_blocks, src/layout.rs:525-533 */
uint32_t wf_effective_pml_thickness_blocks(uint32_t nx, uint32_t ny, uint32_t nz,
                                           uint32_t requested_voxels);
/* PmlProfile1D::build, src/layout.rs:471-513: n entries of b, a, inv_kappa. */
void wf_pml_profile(uint32_t n, float grading_order, float kappa_max, float alpha_max,
                    float target_reflection, float dt, float dx, uint32_t thickness_voxels,
                    float *b, float *a, float *inv_kappa);
/* Waveform::evaluate, src/source.rs:46-59 */
float wf_waveform(int32_t kind, float p0, float p1, float t);
/* amplitude * waveform(t) for t = (float)(first_step+k+1)*dt (src/engine.rs:769),
 * k < nsteps, written as values[k*n + s]. */
int wf_source_values(uint32_t n, const wf_source_desc *src, float dt, uint64_t first_step,
                     uint64_t nsteps, float *values);
/* Probe twiddles for Probe::accumulate (src/probe.rs:126-138): for probe p,
 * frequency i (flattened j = sum_{q<p} nfreq_q + i) and step k:
 * cos_tab[k*nf+j] = cos(2*pi*(double)f * (double)t), sin_tab likewise,
 * active[k*n+p] = !(t < start[p]), with t = (float)(first_step+k+1)*dt. */
int wf_probe_twiddles(uint32_t n, const wf_probe_site *sites, const float *freqs,
                      const float *starts, float dt, uint64_t first_step, uint64_t nsteps,
                      double *cos_tab, double *sin_tab, uint8_t *active);
/* Probe::spectrum, src/probe.rs:143-167 */
void wf_probe_spectrum(uint32_t nfreq, const double *sum_cos, const double *sum_sin,
                       uint64_t n, double *amplitude, double *phase_rad);
/* Scene::parse + Scene::voxelize, src/scene.rs:81-230.  ids (x-fastest,
 * nx*ny*nz, pre-filled by the caller, normally 0) and table (256*4, normally
 * wf_material_table_vacuum) are updated in place.  Returns the number of
 * distinct materials (>= 0) or WF_ERR_INVALID with a message (parse error,
 * more than 255 materials). */
int wf_scene_voxelize(const char *text, uint32_t nx, uint32_t ny, uint32_t nz, float dt,
                      float dx, uint8_t *ids, int32_t *table);
/* voxelize_demo_sphere, src/main.rs:403-425 (material id 1). */
void wf_voxelize_demo_sphere(uint32_t nx, uint32_t ny, uint32_t nz, uint8_t *ids);

/* ---------------------------------------------------------------------- */
/* Device engine                                                            */
/* ---------------------------------------------------------------------- */
/*
 * Creates a device context on HIP device `device` and uploads the scene.
 *   nx, ny, nz      voxels, each a positive multiple of 8 (GridDims::new,
 *                   src/layout.rs:60-67).
 *   material_ids    nx*ny*nz u8, x-fastest (MaterialGrid, src/layout.rs:327);
 *                   NULL = all vacuum (id 0).
 *   coeff_table     256 x {ca, cb, da, db} Q16.16 (MaterialTable); per-voxel
 *                   coefficients are table[id], exactly CoeffGrid::build
 *                   (src/layout.rs:230-260).
 *   pml_shell_blocks NULL = PML disabled (engine::run with pml = None).
 *                   Otherwise [3] per-axis shell depth in blocks: a block is
 *                   in the PML shell iff some axis a has shell[a] > 0 and its
 *                   block coordinate is < shell[a] or >= nb[a]-shell[a]
 *                   (PmlAuxGrid::build, src/layout.rs:594-618).  Requires
 *                   2*shell[a] <= nb[a].
 *   pml_profiles    required with pml_shell_blocks: b, a, inv_kappa for x
 *                   (3*nx floats), then y (3*ny), then z (3*nz) -- the
 *                   PmlProfile1D arrays.  Entries outside [0, 8*shell[a]) and
 *                   [n-8*shell[a], n) must be identity (b=1, a=0, ik=1), as
 *                   PmlProfile1D::build makes them.
 *   periodic_mask   bit a set = axis a wraps instead of reading the zero
 *                   boundary (OUTER_ZERO_BLOCK, src/engine.rs:83); the
 *                   validation studies wrap y and z.
 */
int wf_create(uint32_t nx, uint32_t ny, uint32_t nz, const uint8_t *material_ids,
              const int32_t *coeff_table, const uint32_t *pml_shell_blocks,
              const float *pml_profiles, uint32_t periodic_mask, int device, wf_ctx **out);
void wf_destroy(wf_ctx *ctx);

/* Uploads the source table: n sites and values[k*n + s] for the next nsteps
 * steps (from the current step counter).  Sources are applied in list order
 * after the E update of each step (src/engine.rs:769-772). */
int wf_set_sources(wf_ctx *ctx, uint32_t n, const wf_source_site *sites, uint64_t nsteps,
                   const float *values);
/* The same with shared value columns: source s adds (or assigns)
 * values[k*ncol + cols[s]] at step k, so sources with the same amplitude and
 * waveform share one column (the reference evaluates amplitude *
 * waveform(t) per source per step, src/source.rs:79-95: equal sources give
 * equal values).  A sheet of 1024^2 identical sources over 10^4 steps then
 * needs 40 KB of values instead of 42 GB.  cols = NULL means column s for
 * source s (wf_set_sources, ncol = n). */
int wf_set_source_columns(wf_ctx *ctx, uint32_t n, const wf_source_site *sites,
                          const uint32_t *cols, uint32_t ncol, uint64_t nsteps,
                          const float *values);
/* Uploads the probe table (see wf_probe_twiddles for the layout) and zeroes
 * the accumulators.  trace != 0 also records the sampled f32 value of every
 * probe at every step (read with wf_read_trace). */
int wf_set_probes(wf_ctx *ctx, uint32_t n, const wf_probe_site *sites, uint64_t nsteps,
                  const double *cos_tab, const double *sin_tab, const uint8_t *active,
                  int trace);

/* Advances nsteps full timesteps: H, E, sources, probes per step
 * (src/engine.rs:748-782).  Asynchronous on the context's stream. */
int wf_step(wf_ctx *ctx, uint64_t nsteps);
/* Waits for all work queued on the context. */
int wf_sync(wf_ctx *ctx);
/* Selects the step kernels: on = 0 (default) the exact set, compiled
 * without FP contraction and bit-identical to the CPU engine (the reference
 * never contracts, src/fdtd.rs); on != 0 the same kernels compiled with
 * contraction (a*b+c as fused multiply-adds), within rounding of the exact
 * results.  WF_FP_CONTRACT=1 at wf_create selects the contracted set. */
int wf_set_fp_contract(wf_ctx *ctx, int on);
/* Steps advanced since creation. */
uint64_t wf_step_count(const wf_ctx *ctx);

/* Field state in the reference snapshot format (nx*ny*nz*24 bytes). */
uint64_t wf_field_bytes(const wf_ctx *ctx);
int wf_read_fields(wf_ctx *ctx, void *dst);
int wf_write_fields(wf_ctx *ctx, const void *src);
/* Asynchronous snapshot (the reference's serialize_snapshot into a buffer
 * that is written while the next steps run, src/engine.rs:622-640,
 * 785-793): packs the current state on the device, in stream order after the
 * steps queued so far, and copies it to dst (wf_field_bytes bytes; pinned
 * memory, e.g. hipHostMalloc, lets the copy overlap the following steps)
 * on a side stream.  dst must not be read or reused before wf_wait.  A
 * second call waits on the device for the first copy to finish. */
int wf_snapshot_async(wf_ctx *ctx, void *dst);
/* Waits for the pending wf_snapshot_async copy (the reference's
 * wait-before-reuse of a snapshot buffer). */
int wf_wait(wf_ctx *ctx);

/* The whole timestep loop with snapshot egress; replaces engine::run
 * (src/engine.rs:653-819).  Runs num_steps steps; after the step with loop
 * index k (0-based, counted from this call) it appends a snapshot
 * (wf_field_bytes bytes, the serialize_snapshot layout) to output_path when
 * k % snapshot_every == 0.  The file is created or truncated and written with
 * O_DIRECT where the file system allows it (ordinary writes otherwise), in
 * requests of at most 1 GiB; a failed or short write returns WF_ERR_IO.
 * Snapshots are packed into one of two device buffers and copied out through
 * pinned host chunks by a writer thread while the loop continues; the loop
 * only waits when both buffers are still being written.  *seconds (if not
 * NULL) receives the wall time including the final drain.  Returns after
 * every snapshot is on file.  Source and probe tables must cover the steps. */
int wf_run(wf_ctx *ctx, uint64_t num_steps, uint64_t snapshot_every, const char *output_path,
           double *seconds);

/* ---------------------------------------------------------------------- */
/* Batches of independent scenes                                            */
/* ---------------------------------------------------------------------- */
/* Replaces calling the loop body of engine::run (src/engine.rs:747-782) for
 * each of many small scenes in turn: the scenes (contexts created with
 * wf_create on one device, sources and probes set as usual) advance together,
 * two steps per pass with one two-step launch per CPML variant covering
 * every scene (see wf_batch_two_step), or one H launch and one E launch per
 * step, then one launch each for all sources and all probes.  Each scene's
 * results are bit-identical to stepping it alone.  The batch advances each
 * context's current state (the two-step passes write the context's second
 * state, which becomes the current one) and its step count; contexts stay owned by
 * the caller and must outlive the batch.  Work queued on a context's own
 * stream is waited for at each wf_batch_step; wf_batch_sync before reading a
 * scene back.  A batch runs the exact (uncontracted) kernels whatever a
 * scene's wf_set_fp_contract setting. */
typedef struct wf_batch wf_batch;
int wf_batch_create(uint32_t n, wf_ctx *const *scenes, wf_batch **out);
void wf_batch_destroy(wf_batch *batch);
int wf_batch_step(wf_batch *batch, uint64_t nsteps);
int wf_batch_sync(wf_batch *batch);
/* nsteps steps between two events on the batch stream; *total_ms = their time. */
int wf_batch_time_steps(wf_batch *batch, uint64_t nsteps, double *total_ms);
/* Autotuner of the split steps (a batch on two-step passes takes no steps
 * and reports its current tiling with *steps_taken = 0): runs
 * steps_per_candidate real steps (the first untimed) under
 * each candidate tiling and schedule (z-chunk, lanes-per-row cap -- a power
 * of two up to 64, or 0 for rows fitted to each scene's width --, grid: 0 =
 * one workgroup per tile, n = n persistent workgroups striding over the
 * tiles), keeps the fastest and reports it; *steps_taken = steps advanced
 * (results do not depend on the tiling).  WF_BATCH_ZCHUNK sets the initial
 * z-chunk (default 8). */
int wf_batch_tune(wf_batch *batch, uint32_t steps_per_candidate, int *zchunk,
                  int *lanes_per_row, int *grid, uint64_t *steps_taken);
int wf_batch_set_tiling(wf_batch *batch, int zchunk, int lanes_per_row, int grid);
/* Whether the batch advances two steps per pass (*two_step = 1: every scene
 * fused, no periodic axis, WF_BATCH_TWOSTEP not 0; each scene's regions --
 * the single-scene split into CPML shell and interior ranges, or the whole
 * scene in one box when that split does not fit -- swept by one launch per
 * CPML variant over all scenes, an odd last step on the H/E kernels above)
 * and the algorithmic HBM bytes of one pass (49 B per cell, 32 B per graded
 * axis of a cell). */
int wf_batch_two_step(const wf_batch *batch, int *two_step, uint64_t *pass_bytes);
/* Sum of the scenes' cells. */
uint64_t wf_batch_cells(const wf_batch *batch);

/* ---------------------------------------------------------------------- */
/* Learned thin absorbing boundary                                          */
/* ---------------------------------------------------------------------- */
/* No counterpart in the reference, which terminates every face with a
 * graded CPML backed by the zero field (OUTER_ZERO_BLOCK, src/engine.rs:83;
 * PmlProfile1D, src/layout.rs:471-513).  Here a face can instead end in a
 * thin layer whose outside "ghost" fields (the zero the boundary cell's curl
 * update reads) are predicted each half step by a small per-face MLP:
 * 16 inputs (4-cell-deep E/H/psi stencil and a short history, scaled by
 * their largest magnitude) -> 64 -> 64 -> 1, hardtanh, BF16 weights and
 * activations on MFMA with FP32 accumulation, plus an FP32 linear skip.
 *
 * Parameters of one network: wf_mlp_param_count() floats, laid out
 * W1[16][64] b1[64] W2[64][64] b2[64] W3[64] b3 wlin[16] (W[in][out]).
 * Faces: 0 x_lo, 1 x_hi, 2 y_lo, 3 y_hi, 4 z_lo, 5 z_hi. */
#define WF_MLP_INPUTS 16
#define WF_SAMPLE_STRIDE 20 /* recorded sample: 16 inputs, target, 3 pad */
uint32_t wf_mlp_param_count(void);

/* Training on the GPU: full-batch Adam on the mean squared error of
 * n samples (inputs[n][16], targets[n], both already scaled). */
typedef struct wf_mlp_trainer wf_mlp_trainer;
int wf_mlp_trainer_create(int device, uint64_t n, const float *inputs, const float *targets,
                          wf_mlp_trainer **out);
void wf_mlp_trainer_destroy(wf_mlp_trainer *t);
/* Sets the FP32 master parameters (and resets the Adam moments). */
int wf_mlp_set_params(wf_mlp_trainer *t, const float *params);
/* Output head: WF_HEAD_DIRECT (default) -- the network output is the scaled
 * ghost value; WF_HEAD_ONE_WAY -- the output y sets the coefficient of a
 * first-order one-way update, ghost/s = P0(n-1) + tanh(y) (P0(n) - g(n-1)),
 * from sign-normalized inputs (a contractive recursion on the ghost). */
#define WF_HEAD_DIRECT 0
#define WF_HEAD_ONE_WAY 1
/* WF_HEAD_DELTA -- the linear skip plus the network's output (on
 * sign-normalized inputs) times the boundary field's step, ghost/s =
 * skip(u) + m (P0(n) - P0(n-1)): the network adds nothing at zero frequency. */
#define WF_HEAD_DELTA 2
/* WF_HEAD_HIGDON -- the second-order Higdon update, a product of two
 * first-order one-way factors (1 + k g T) - X (g T + k) (T: one step back,
 * X: one cell inward, g <= 1 damps the factor's memory): a fixed one with
 * k = wlin[0], g = wlin[1] and an adaptive one with k = tanh(y), y the
 * network's output on sign-normalized inputs, g = wlin[2]; 0 < wlin[3] < 1
 * smooths the adaptive k in time at each face position (an exponential
 * average with weight wlin[3] of tanh(y) - tanh(b3), added to tanh(b3));
 * the other skip weights are not used.  P at depth 1 and its history
 * (inputs 1, 10, 11) enter besides P0 and the ghost's history.  With the
 * network's output weights zero this is the fixed second-order Higdon
 * update. */
#define WF_HEAD_HIGDON 4
/* WF_HEAD_DIFF -- ghost/s = P0(n) + skip(d) + m(d), network and skip on
 * differences of the inputs that vanish for a field constant in space and
 * time (depth-to-depth steps of P and Q, inputs 8, 9 as they are,
 * P1(n-1) - P1, P1(n-2) - P1(n-1), g(n-1) - P0, g(n-2) - g(n-1),
 * P0(n-1) - P0, P0(n-2) - P0(n-1)): a static
 * uniform field is held exactly whatever the weights. */
#define WF_HEAD_DIFF 3
int wf_mlp_set_head(wf_mlp_trainer *t, int head);
int wf_mlp_get_params(wf_mlp_trainer *t, float *params);
/* Gradient of the loss at the current parameters (the network as it runs:
 * BF16 weights and activations) and the loss itself. */
int wf_mlp_gradient(wf_mlp_trainer *t, float *grad, double *loss);
/* iters Adam steps at learning rate lr (lr_skip for the linear skip wlin;
 * 0 keeps it fixed); odd != 0 leaves the biases unchanged (zero biases make
 * the network odd in its inputs); *loss = loss after the last. */
int wf_mlp_train(wf_mlp_trainer *t, uint32_t iters, float lr, float lr_skip, int odd,
                 double *loss);
/* Network output for n scaled input rows. */
int wf_mlp_predict(wf_mlp_trainer *t, uint64_t n, const float *inputs, float *outputs);
/* Allocation-free wf_mlp_predict.  No counterpart in the reference; for
 * callers that cannot allocate device memory (tight evaluation loops,
 * harnesses that forbid hipMalloc).  Evaluates the network `params`
 * (wf_mlp_param_count() floats) with output head `head` on n scaled input
 * rows (inputs[n][16] -> outputs[n], host memory) on HIP device `device`,
 * with the same kernels and the same results as wf_mlp_predict, staging
 * parameters, packed network, inputs and outputs in a static device array
 * that each device holds until the process ends: no hipMalloc, hipFree or
 * stream creation.  n at most wf_mlp_predict_noalloc_max() (65536), else
 * WF_ERR_INVALID before any device work.  Returns after the outputs are
 * written.  All calls on one device share its array: calls that overlap on
 * one device (from several threads) race and must be serialized by the
 * caller.  Work is queued on the device's null stream. */
int wf_mlp_predict_noalloc(int device, const float *params, int head, uint64_t n,
                           const float *inputs, float *outputs);
uint64_t wf_mlp_predict_noalloc_max(void);

/* Predicted faces: the faces in face_mask get ghost planes predicted by the
 * networks params[f * wf_mlp_param_count() ...] with output head `head` each
 * half step (high faces before the H phase, low faces before the E phase);
 * other faces keep their
 * boundary.  A face's axis must not be periodic.  Clears the faces' history.
 * face_mask 0 removes nothing; wf_clear_faces removes every hook. */
int wf_set_learned_faces(wf_ctx *ctx, uint32_t face_mask, const float *params, int head);
/* Limits on the predicted ghost values (defaults: passive = 1, retain = 0.5,
 * gamma = 8).  passive != 0: the face keeps an energy credit C, the discrete
 * Yee energy that has left the domain through its ghosts so far (per ghost
 * and step sigma g (Q^a + Q^b) / 2, g the ghost value, Q the boundary field
 * its curl term updates, before and after that update, sigma = +1 / -1 for
 * polarization 0 / 1), and A, the energy absorbed in steps of net outflow;
 * every ghost value may draw an equal share of C - retain A in a step, a
 * prediction beyond it moves to the nearest value within it, and ghosts
 * read by one boundary cell on an edge of two limited faces are limited as
 * a pair.  The face then never returns more than (1 - retain) of the energy
 * it has absorbed, so the fields stay bounded whatever the network
 * predicts (the energy of the staggered Yee scheme is not positive for a
 * ghost that grows on its own; gamma > 0 bounds |ghost| by gamma times the
 * largest |P| of its 4-cell stencil, which rules that out).  retain in
 * [0, 1).  Applies to the predicted faces now set and later ones. */
int wf_set_face_limits(wf_ctx *ctx, int passive, float retain, float gamma);
/* The same limits on replayed faces (wf_replay_faces; default off): the
 * recorded ghosts of a thick-CPML run take the place of the predictions, a
 * check of the limits against the exact boundary. */
int wf_set_replay_limits(wf_ctx *ctx, int passive, float retain, float gamma);
/* State of a predicted (or limited replayed) face: 2 (polarizations) x
 * positions x 8 floats: g(n-1), g(n-2), P0(n-1), P0(n-2), the last step's
 * flux through the ghost, steps limited by the energy bound, steps limited by
 * the magnitude bound, sum of squared corrections.  Positions run fast axis
 * first (the ghost plane's order). */
int wf_read_face_state(wf_ctx *ctx, uint32_t face, float *dst);
int wf_clear_faces(wf_ctx *ctx);

/* Recording faces (training data from a thick-CPML run): a virtual face of
 * the thin domain placed inside this domain, b0 = normal coordinate of its
 * boundary cell, the ghost cell one further out; fast/slow ranges on the
 * lower-/higher-numbered tangential axis.  From the next step on, every
 * `every` steps (up to `rows` times) each position and polarization writes
 * a sample (the same 16 inputs a predicted face computes, and the true ghost
 * value / s); the history uses the true ghost values. */
typedef struct wf_face_region {
    int32_t b0, fast0, nfast, slow0, nslow;
} wf_face_region;
int wf_record_face_samples(wf_ctx *ctx, uint32_t face_mask, const wf_face_region *regions,
                           uint64_t every, uint64_t rows);
/* Replayed faces: ghost planes taken step by step from samples recorded
 * by a thick-CPML run (samples[f]: rows * 2 * positions * WF_SAMPLE_STRIDE
 * floats as wf_read_face_samples returns them, recorded every step over the
 * matching region) -- the perfect predictor, for checks and bounds. */
int wf_replay_faces(wf_ctx *ctx, uint32_t face_mask, uint64_t rows, const float *const *samples);
/* Samples a recording face holds (rows * 2 * positions), and a copy of them
 * (count * WF_SAMPLE_STRIDE floats: 16 scaled inputs, target / s, target, s,
 * padding; unwritten rows are zero). */
uint64_t wf_face_sample_count(const wf_ctx *ctx, uint32_t face);
int wf_read_face_samples(wf_ctx *ctx, uint32_t face, float *dst);
/* CPML psi memory in the reference PmlAux layout (src/layout.rs:545-558):
 * 12 x 512 f32 per shell block, shell blocks in block-major order. */
uint64_t wf_psi_shell_blocks(const wf_ctx *ctx);
int wf_read_psi(wf_ctx *ctx, float *dst);

/* Probe accumulators: flattened over (probe, frequency) like the twiddles;
 * n[p] = samples recorded by probe p. */
int wf_read_probe_sums(wf_ctx *ctx, double *sum_cos, double *sum_sin, uint64_t *n);
/* trace[k*nprobes + p] for the steps since wf_set_probes (k < steps taken). */
int wf_read_trace(wf_ctx *ctx, float *trace, uint64_t max_steps);

/* ---------------------------------------------------------------------- */
/* Measurement                                                              */
/* ---------------------------------------------------------------------- */
/* Times nsteps full steps with a HIP event pair on the context's stream;
 * writes the elapsed milliseconds. */
int wf_time_steps(wf_ctx *ctx, uint64_t nsteps, double *total_ms);
/* Runs nsteps full steps with a HIP event pair around every step-kernel
 * launch (same stream); writes the mean per-launch milliseconds of three
 * kernel groups, by wf_kernel_variant: "fused" -- the single-pass kernel
 * over the interior box, the H-phase and the E-phase kernels over the frame
 * of cells around it (which holds the CPML); "split" -- 0, the H and the E
 * kernel over the whole domain; "twostep" -- per pass of two steps, the
 * two-step kernels over the interior region, over the x shell regions and
 * over the y and z shell regions; "twostep-rows" -- per pass, the batched
 * two-step launches together in fused_ms, 0, 0. */
int wf_time_kernels(wf_ctx *ctx, uint64_t nsteps, double *fused_ms, double *h_ms,
                    double *e_ms);
/* Algorithmic HBM bytes of one launch of each group above.  Fused kernel:
 * 24 B read + 24 B written + 1 B material id per interior cell.  H or E
 * kernel: per cell it covers 12 B read of the curl field, 12 B read + 12 B
 * write of the updated field and 1 B id, plus 8 B (read + write) per
 * non-identity psi term of that phase.  Two-step kernels: 49 B per cell for
 * the two steps and 32 B per graded axis of a cell. */
int wf_kernel_bytes(const wf_ctx *ctx, uint64_t *fused_bytes, uint64_t *h_bytes,
                    uint64_t *e_bytes);
/* The step scheme in use: "twostep" (two steps per pass on the two-step
 * kernels over the CPML regions and the interior, state A -> B), "twostep-
 * rows" (two steps per pass through the batched two-step kernel over full
 * x rows), "fused" (one step per pass: the single-pass kernel over the
 * interior, split H/E kernels over the CPML frame, ping-pong state) or
 * "split" (H kernel then E kernel over the whole domain, in place; chosen
 * at wf_create by WF_KERNEL=split). */
const char *wf_kernel_variant(const wf_ctx *ctx);
/* Schedule chosen at wf_create: 0 two-step regions, 1 two-step rows, 2
 * single steps (WF_SCHEDULE sets it; otherwise grids under 2^25 cells time
 * the three on the zeroed state and keep the fastest, larger grids take 0).
 * probe_ms (may be null): the three timings in ms per step, -1 where not
 * timed.  wf_kernel_variant reports what runs (face hooks, periodic axes or
 * the contracted set fall back to other schemes).  No counterpart in the
 * reference. */
int wf_schedule(const wf_ctx *ctx, int *schedule, double *probe_ms);
/* Field placement chosen by wf_create: the step time (ms) of each
 * arrangement it timed, in order (up to `cap` written; the initial buffers
 * first, then each extra state buffer in place of state A and, in the fused
 * variant, of state B), their number in *n (0 when no probe ran) and the
 * index of the one kept in *kept.  No counterpart in the reference: on
 * MI355X the step kernels over one allocation of a 24 GiB state buffer run
 * up to ~12% faster or slower than over another, depending on where the
 * driver placed its pages, so for grids whose fields take >= 8 GiB per state
 * wf_create tries eight extra state buffers (one held at a time, four
 * steps timed per arrangement: about 1 s at 1024^3; no new buffer after
 * 4 s) and keeps the fastest arrangement (WF_PLACEMENT=n tries n, 0 = off,
 * at most 8; WF_PLACEMENT_MIN_MIB moves the threshold). */
int wf_placement(const wf_ctx *ctx, double *ms, int cap, int *n, int *kept);
/* Copies `bytes` (read + write = 2*bytes of HBM traffic) iters times with a
 * plain 16-byte-per-lane copy kernel on `device`; writes the mean ms per
 * copy.  No counterpart in the reference: the measured HBM ceiling and the
 * scale for the FETCH_SIZE / WRITE_SIZE counters. */
int wf_hbm_copy(uint64_t bytes, int device, int iters, double *ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif

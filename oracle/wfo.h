/*
 * wfo.h -- CPU oracle for the wavefront FDTD timestep path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * Rust engine (AndrejRumenovski/wavefront), kept in the reference's own
 * block-major AoSoA layout so that its field bytes *are* the reference
 * snapshot format.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker / CPU baseline.
 * The product (wavefront_amd/) never links or calls it.
 *
 * Every arithmetic expression follows the reference's evaluation order and
 * is compiled with -ffp-contract=off (no FMA contraction, no fast-math);
 * libm-dependent values come from glibc (expf, logf, powf, sinf, asinf,
 * sin, cos, atan2), which is what Rust's f32/f64 methods lower to.
 *
 * Pinning (how each function is checked against something other than itself):
 *   - q16 / material coeffs / update kernels (plain) / sinusoid waveform /
 *     f64 I-Q demodulation: validation/convergence_data.csv (all 4 rows,
 *     bit-exact) via wfo_convergence_study.
 *   - PML profile (expf/powf/logf), X-axis CPML kernels, two-run reflection:
 *     validation/pml_reflection_data.csv (all 4 rows, bit-exact) via
 *     wfo_pml_reflection_study.
 *   - 3-D block indexing, zero boundary, Y/Z CPML axes, edge/corner shell
 *     blocks: cross-checked against an independent brute-force scalar
 *     restatement (tests/test_oracle_invariants.py) and against symmetry /
 *     decomposition-invariance properties.
 *   - Gaussian / Ricker waveforms, probe spectrum: restatement plus the
 *     reference's own tolerance tests (src/source.rs:103-129,
 *     src/probe.rs:188-260) -- bit-level parity with the Rust binary is
 *     unpinned for these (glibc expf is shared by oracle and product).
 */
#ifndef WFO_H
#define WFO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WFO_BLOCK_DIM 8
#define WFO_VOX_PER_BLOCK 512

/* src/layout.rs:114-125 */
int32_t wfo_q16_from_f32(float v);
float wfo_q16_to_f32(int32_t q);

/* src/layout.rs:194-208; out = {ca, cb, da, db} as Q16.16 */
void wfo_material_coeffs(float eps_r, float mu_r, float sigma, float dt, float d,
                         int32_t out[4]);

/* src/layout.rs:525-533 */
size_t wfo_effective_pml_thickness_blocks(size_t nx, size_t ny, size_t nz,
                                          size_t requested_voxels);

/* src/layout.rs:471-513.  Writes n entries to b, a, ik. */
void wfo_pml_profile(size_t n, float grading_order, float kappa_max, float alpha_max,
                     float target_reflection, float dt, float dx,
                     size_t thickness_voxels, float *b, float *a, float *ik);

/* src/source.rs:46-59.  kind 0 = Gaussian(p0=t0, p1=spread),
 * 1 = Sinusoid(p0=freq), 2 = Ricker(p0=peak_freq, p1=t0). */
float wfo_waveform(int kind, float p0, float p1, float t);

/* ---- simulation state (FieldGrid + CoeffGrid + PmlAuxGrid) ---------- */
typedef struct wfo_sim wfo_sim;

/*
 * ids: nx*ny*nz material ids, x-fastest (MaterialGrid, src/layout.rs:327),
 *      NULL = all vacuum (id 0).
 * table: 256 x {ca,cb,da,db} Q16.16 (MaterialTable).
 * shell_blocks[3]: per-axis PML shell depth in blocks (0 = that axis does
 *      not put blocks in the shell).  The engine uses the same value on all
 *      axes; the reflection study uses x only.
 * profiles: NULL (PML disabled: plain kernels everywhere), or b,a,ik arrays
 *      for x (3*nx floats) then y (3*ny) then z (3*nz).
 * periodic_mask: bit0/1/2 = wrap x/y/z neighbours instead of the zero
 *      boundary (the studies wrap y and z, examples/convergence_study.rs:121).
 */
wfo_sim *wfo_sim_create(size_t nx, size_t ny, size_t nz, const uint8_t *ids,
                        const int32_t *table, const size_t shell_blocks[3],
                        const float *profiles, unsigned periodic_mask);
void wfo_sim_destroy(wfo_sim *s);

/* comp 0..2 = Ex,Ey,Ez.  hard != 0 assigns instead of adding (the studies'
 * sheet source). */
int wfo_sim_add_source(wfo_sim *s, size_t x, size_t y, size_t z, int comp, int hard,
                       float amplitude, int kind, float p0, float p1);
/* comp 0..5 = Ex..Hz */
int wfo_sim_add_probe(wfo_sim *s, size_t x, size_t y, size_t z, int comp,
                      size_t nfreq, const float *freqs, float start);

/* Runs steps [first_step, first_step+nsteps) exactly like engine::run's loop
 * (src/engine.rs:747-782).  trace (optional): nsteps x nprobes f32 samples,
 * read after the sources, recorded every step regardless of the probe's
 * start time.  threads = slab workers (src/engine.rs:666-669). */
void wfo_sim_run(wfo_sim *s, size_t first_step, size_t nsteps, float dt, int threads,
                 float *trace);

/* The timestep without sources/probes (step_all_slabs). */
void wfo_sim_step_fields(wfo_sim *s, int threads);

/* Probe::accumulate(&grid, t) of every probe on the current fields
 * (src/probe.rs:126-138), outside the step loop. */
void wfo_sim_accumulate_probes(wfo_sim *s, float t);

/* Snapshot bytes in the reference format (src/engine.rs:622-640). */
size_t wfo_sim_bytes(const wfo_sim *s);
void wfo_sim_read_aosoa(const wfo_sim *s, void *dst);
void wfo_sim_write_aosoa(wfo_sim *s, const void *src);
/* psi memory: 12 arrays x 512 per shell block, in the PmlAux field order
 * (src/layout.rs:545-558); blocks in block-major order, shell blocks only. */
size_t wfo_sim_shell_block_count(const wfo_sim *s);
void wfo_sim_read_psi(const wfo_sim *s, float *dst);

size_t wfo_sim_probe_count(const wfo_sim *s);
size_t wfo_sim_probe_nfreq(const wfo_sim *s, size_t p);
void wfo_sim_probe_sums(const wfo_sim *s, size_t p, double *sum_cos, double *sum_sin,
                        uint64_t *n);
/* src/probe.rs:143-167 */
void wfo_probe_spectrum(size_t nfreq, const double *sum_cos, const double *sum_sin,
                        uint64_t n, double *amplitude, double *phase);

/* ---- scenes (src/scene.rs) ------------------------------------------ */
/* Voxelizes scene text into ids (x-fastest, caller-zeroed) and the table.
 * Returns the number of distinct materials, or -1 on parse error / >255
 * materials (msg gets a reason). */
int wfo_scene_voxelize(const char *text, size_t nx, size_t ny, size_t nz, float dt,
                       float dx, uint8_t *ids, int32_t *table, char *msg, size_t msg_len);
/* src/main.rs:403-425 (id 1) */
void wfo_voxelize_demo_sphere(size_t nx, size_t ny, size_t nz, uint8_t *ids);

/* ---- validation studies --------------------------------------------- */
/* examples/convergence_study.rs: out[r*6 + c] for the 4 rows and 6 CSV
 * columns, each an f32 value. */
void wfo_convergence_study(float *out, int threads);
/* examples/pml_reflection_study.rs: 4 reflection coefficients. */
void wfo_pml_reflection_study(double *out, int threads);

/* examples/convergence_study.rs:67-72 */
float wfo_yee_phase_velocity(float dx, float dt, float freq_hz);

#ifdef __cplusplus
}
#endif
#endif

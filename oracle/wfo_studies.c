/*
 * wfo_studies.c -- restatement of the reference's two validation studies,
 * examples/convergence_study.rs and examples/pml_reflection_study.rs, on the
 * oracle engine.  Their outputs are the reference-held golden vectors
 * validation/convergence_data.csv and validation/pml_reflection_data.csv.
 *
 * TEST INFRASTRUCTURE ONLY (see wfo.h).
 *
 * The studies' own serial loops (convergence_study.rs:110-176,
 * pml_reflection_study.rs:135-195) are the engine's Jacobi H-then-E step with
 * Y/Z wrapped periodically, followed by a hard sheet source and a probe read;
 * wfo_sim with periodic_mask = y|z, hard sources and a trace reproduces them
 * operation for operation.
 */
#define _GNU_SOURCE
#include "wfo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* examples/convergence_study.rs:62 */
static const float C0 = 299792458.0f;
static const float PI_F = 3.14159265358979323846f;

static size_t nx8(size_t v) { return (v + WFO_BLOCK_DIM - 1) / WFO_BLOCK_DIM * WFO_BLOCK_DIM; }

static size_t sat_usize(float r)
{
    if (isnan(r) || r <= 0.0f) return 0;
    return (size_t)r;
}

/* examples/convergence_study.rs:67-72 */
float wfo_yee_phase_velocity(float dx, float dt, float freq_hz)
{
    float omega = 2.0f * PI_F * freq_hz;
    float rhs = (dx / (C0 * dt)) * sinf(omega * dt / 2.0f);
    float k = (2.0f / dx) * asinf(rhs);
    return omega / k;
}

/* run_plane_wave: convergence_study.rs:81-179 / pml_reflection_study.rs:85-198.
 * pml_thickness_voxels == 0 -> plain kernels, zero X boundary. */
static float *run_plane_wave(size_t nx, float dx, float dt, float freq_hz, size_t x_source,
                             size_t x_probe, size_t steps, size_t pml_thickness_voxels,
                             int threads)
{
    int32_t table[256 * 4];
    int32_t vac[4];
    wfo_material_coeffs(1.0f, 1.0f, 0.0f, dt, dx, vac);
    for (int i = 0; i < 256; i++) memcpy(table + 4 * i, vac, sizeof(vac));

    size_t nbx = nx / WFO_BLOCK_DIM;
    size_t shell[3] = {0, 0, 0};
    float *prof = NULL;
    if (pml_thickness_voxels > 0) {
        /* pml_reflection_study.rs:116-126: X-only profile and shell */
        size_t tb = (pml_thickness_voxels + WFO_BLOCK_DIM - 1) / WFO_BLOCK_DIM;
        if (tb > nbx / 2) tb = nbx / 2;
        size_t tv = tb * WFO_BLOCK_DIM;
        size_t n[3] = {nx, WFO_BLOCK_DIM, WFO_BLOCK_DIM};
        prof = malloc(3 * (n[0] + n[1] + n[2]) * sizeof(float));
        float *p = prof;
        for (int ax = 0; ax < 3; ax++) {
            /* PmlConfig::default(), src/layout.rs:425-435 */
            wfo_pml_profile(n[ax], 3.0f, 7.0f, 0.05f, 1.0e-6f, dt, dx, ax == 0 ? tv : 0, p,
                            p + n[ax], p + 2 * n[ax]);
            p += 3 * n[ax];
        }
        shell[0] = tb;
    }
    wfo_sim *s = wfo_sim_create(nx, WFO_BLOCK_DIM, WFO_BLOCK_DIM, NULL, table, shell, prof,
                                6u /* wrap y and z */);
    free(prof);
    for (size_t ly = 0; ly < WFO_BLOCK_DIM; ly++)
        for (size_t lz = 0; lz < WFO_BLOCK_DIM; lz++)
            wfo_sim_add_source(s, x_source, ly, lz, 2, 1, 1.0f, 1, freq_hz, 0.0f);
    wfo_sim_add_probe(s, x_probe, 0, 0, 2, 0, NULL, 0.0f);
    float *trace = malloc(steps * sizeof(float));
    wfo_sim_run(s, 0, steps, dt, threads, trace);
    wfo_sim_destroy(s);
    return trace;
}

/* convergence_study.rs:255-282 / pml_reflection_study.rs:204-224 */
static void quadrature(const float *trace, size_t steps, float dt, double omega,
                       float skip_before_s, double *a_out, double *b_out)
{
    double sum_cos = 0.0, sum_sin = 0.0;
    uint32_t count = 0;
    for (size_t i = 0; i < steps; i++) {
        float t = ((float)i + 1.0f) * dt;
        if (t > skip_before_s) {
            double phase = omega * (double)t;
            sum_cos += (double)trace[i] * cos(phase);
            sum_sin += (double)trace[i] * sin(phase);
            count++;
        }
    }
    *a_out = 2.0 * sum_cos / (double)count;
    *b_out = 2.0 * sum_sin / (double)count;
}

/* examples/convergence_study.rs:204-311 */
void wfo_convergence_study(float *out, int threads)
{
    float freq_hz = 3.0e10f;
    float wavelength_m = C0 / freq_hz;
    float courant = 0.4f;
    float periods_traveled = 8.0f;
    const float ppws[4] = {10.0f, 15.0f, 20.0f, 30.0f};

    for (int r = 0; r < 4; r++) {
        float ppw = ppws[r];
        float dx = wavelength_m / ppw;
        float dt = courant * dx / C0;
        float period_s = 1.0f / freq_hz;
        float period_steps = ppw / courant;

        size_t n_voxels = sat_usize(roundf(periods_traveled * ppw));
        size_t x_source = 3 * n_voxels;
        size_t x_probe = x_source + n_voxels;
        size_t nx = nx8(7 * n_voxels);
        size_t steps = sat_usize(ceilf((periods_traveled + 12.0f) * period_steps));
        float expected_delay_s = ((float)n_voxels * dx) / C0;
        double omega = 2.0 * M_PI * (double)freq_hz;

        float *trace = run_plane_wave(nx, dx, dt, freq_hz, x_source, x_probe, steps, 0,
                                      threads);
        float skip_before_s = expected_delay_s + 5.0f * period_s;
        double a, b;
        quadrature(trace, steps, dt, omega, skip_before_s, &a, &b);
        free(trace);

        double raw_delay = atan2(-a, b) / omega;
        double k = round(((double)expected_delay_s - raw_delay) / (double)period_s);
        float measured_delay_s = (float)(raw_delay + k * (double)period_s);
        float measured_v = ((float)n_voxels * dx) / measured_delay_s;
        float theory_v = wfo_yee_phase_velocity(dx, dt, freq_hz);

        float *row = out + 6 * r;
        row[0] = dx;
        row[1] = ppw;
        row[2] = measured_v;
        row[3] = theory_v;
        row[4] = fabsf(measured_v - C0) / C0;
        row[5] = fabsf(theory_v - C0) / C0;
    }
}

/* examples/pml_reflection_study.rs:237-320 */
void wfo_pml_reflection_study(double *out, int threads)
{
    float freq_hz = 3.0e10f;
    float wavelength_m = C0 / freq_hz;
    float courant = 0.4f;
    float ppw = 20.0f;
    float dx = wavelength_m / ppw;
    float dt = courant * dx / C0;
    float period_s = 1.0f / freq_hz;

    size_t probe_offset_voxels = sat_usize(roundf(3.0f * ppw));
    size_t right_gap_voxels = sat_usize(roundf(2.0f * ppw));
    const size_t tbs[4] = {1, 2, 4, 8};

    for (int r = 0; r < 4; r++) {
        size_t thickness_voxels = tbs[r] * WFO_BLOCK_DIM;

        size_t x_source_test = thickness_voxels + sat_usize(roundf(2.0f * ppw));
        size_t x_probe_test = x_source_test + probe_offset_voxels;
        size_t x_pml_face_test = x_probe_test + right_gap_voxels;
        size_t nx_test = nx8(x_pml_face_test + thickness_voxels);

        float incident_delay_s = (float)probe_offset_voxels * dx / C0;
        size_t round_trip_voxels = 2 * (x_pml_face_test - x_probe_test);
        float round_trip_s = (float)round_trip_voxels * dx / C0;
        float settle_periods = 6.0f;
        float skip_before_s = incident_delay_s + round_trip_s + settle_periods * period_s;
        float measurement_periods = 10.0f;
        size_t steps =
            sat_usize(ceilf((skip_before_s + measurement_periods * period_s) / dt)) + 4;

        float *test = run_plane_wave(nx_test, dx, dt, freq_hz, x_source_test, x_probe_test,
                                     steps, thickness_voxels, threads);

        size_t required =
            sat_usize(ceilf(((float)steps * dt + 2.0f * period_s) * C0 / dx)) + WFO_BLOCK_DIM;
        size_t x_probe_ref = required + probe_offset_voxels;
        size_t x_source_ref = x_probe_ref - probe_offset_voxels;
        size_t nx_ref = nx8(x_probe_ref + required);
        float *ref = run_plane_wave(nx_ref, dx, dt, freq_hz, x_source_ref, x_probe_ref, steps,
                                    0, threads);

        double omega = 2.0 * M_PI * (double)freq_hz;
        double a, b;
        quadrature(ref, steps, dt, omega, skip_before_s, &a, &b);
        double incident_amp = sqrt(a * a + b * b);
        float *diff = malloc(steps * sizeof(float));
        for (size_t i = 0; i < steps; i++) diff[i] = test[i] - ref[i];
        quadrature(diff, steps, dt, omega, skip_before_s, &a, &b);
        double reflected_amp = sqrt(a * a + b * b);
        out[r] = reflected_amp / incident_amp;
        free(diff);
        free(test);
        free(ref);
    }
}

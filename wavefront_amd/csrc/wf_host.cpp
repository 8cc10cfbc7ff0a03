// Host-side table builders: the per-run constants the reference computes on
// the CPU before (or alongside) its timestep loop.  Every f32 expression
// keeps the reference's evaluation order; the library is built with
// -ffp-contract=off so each operation rounds separately, and transcendental
// values come from the C library (expf, logf, powf, sinf, cos, sin), which is
// what the reference's f32/f64 methods call.
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <strings.h>
#include <unordered_map>
#include <vector>

#include "wf_internal.h"

namespace wf {

static thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() { g_last_error.clear(); }

bool fault_injected(const char *name)
{
    const char *v = std::getenv("WF_FAULT_INJECT");
    return v && std::strcmp(v, name) == 0;
}

// Rust `f32 as i32` / `f32 as usize`: saturating, NaN -> 0.
static int32_t sat_i32(float r)
{
    if (std::isnan(r)) return 0;
    if (r >= 2147483648.0f) return INT32_MAX;
    if (r <= -2147483648.0f) return INT32_MIN;
    return static_cast<int32_t>(r);
}

static size_t sat_usize(float r)
{
    if (std::isnan(r) || r <= 0.0f) return 0;
    if (r >= 18446744073709551616.0f) return SIZE_MAX;
    return static_cast<size_t>(r);
}

}  // namespace wf

using namespace wf;

extern "C" {

const char *wf_last_error(void) { return g_last_error.c_str(); }

// src/layout.rs:114-116
int32_t wf_q16_from_f32(float v) { return sat_i32(std::roundf(v * 65536.0f)); }

// src/layout.rs:123-125
float wf_q16_to_f32(int32_t q) { return static_cast<float>(q) * (1.0f / 65536.0f); }

// src/layout.rs:194-208
void wf_material_coeffs(float eps_r, float mu_r, float sigma, float dt, float d,
                        int32_t out[4])
{
    const float eps = eps_r * kEps0;
    const float mu = mu_r * kMu0;
    const float loss = sigma * dt / (2.0f * eps);
    const float ca = (1.0f - loss) / (1.0f + loss);
    const float cb = (dt / (eps * d)) / (1.0f + loss);
    const float da = 1.0f;
    const float db = dt / (mu * d);
    out[0] = wf_q16_from_f32(ca);
    out[1] = wf_q16_from_f32(cb);
    out[2] = wf_q16_from_f32(da);
    out[3] = wf_q16_from_f32(db);
}

// src/layout.rs:175-180
void wf_material_table_vacuum(float dt, float d, int32_t *table)
{
    int32_t v[4];
    wf_material_coeffs(1.0f, 1.0f, 0.0f, dt, d, v);
    for (int i = 0; i < 256; ++i) std::memcpy(table + 4 * i, v, sizeof(v));
}

// src/layout.rs:525-533
uint32_t wf_effective_pml_thickness_blocks(uint32_t nx, uint32_t ny, uint32_t nz,
                                           uint32_t requested_voxels)
{
    if (requested_voxels == 0) return 0;
    const uint32_t nbx = nx / WF_BLOCK_DIM, nby = ny / WF_BLOCK_DIM, nbz = nz / WF_BLOCK_DIM;
    // div_ceil without the overflow of (requested + 7) near UINT32_MAX
    uint32_t req = requested_voxels / WF_BLOCK_DIM + (requested_voxels % WF_BLOCK_DIM != 0);
    if (req < 1) req = 1;
    uint32_t lim = nbx < nby ? nbx : nby;
    lim = (lim < nbz ? lim : nbz) / 2;
    if (lim < 1) lim = 1;
    return req < lim ? req : lim;
}

// src/layout.rs:471-513
void wf_pml_profile(uint32_t n, float grading_order, float kappa_max, float alpha_max,
                    float target_reflection, float dt, float dx, uint32_t thickness,
                    float *b, float *a, float *inv_kappa)
{
    for (uint32_t i = 0; i < n; ++i) {
        b[i] = 1.0f;
        a[i] = 0.0f;
        inv_kappa[i] = 1.0f;
    }
    if (thickness == 0) return;
    const float depth = static_cast<float>(thickness) * dx;
    const float eta0 = std::sqrt(kMu0 / kEps0);
    const float sigma_max =
        -(grading_order + 1.0f) * std::log(target_reflection) / (2.0f * eta0 * depth);
    for (uint32_t g = 0; g < thickness; ++g) {
        const float rho = static_cast<float>(thickness - g) / static_cast<float>(thickness);
        const float sigma = sigma_max * std::pow(rho, grading_order);
        const float kappa = 1.0f + (kappa_max - 1.0f) * std::pow(rho, grading_order);
        const float alpha = alpha_max * (1.0f - rho);
        const float bb = std::exp(-(sigma / kappa + alpha) * dt / kEps0);
        const float denom = kappa * (sigma + kappa * alpha);
        const float aa = std::fabs(denom) > 1.0e-30f ? sigma * (bb - 1.0f) / denom : 0.0f;
        const float ik = 1.0f / kappa;
        b[g] = b[n - 1 - g] = bb;
        a[g] = a[n - 1 - g] = aa;
        inv_kappa[g] = inv_kappa[n - 1 - g] = ik;
    }
}

// src/source.rs:46-59
float wf_waveform(int32_t kind, float p0, float p1, float t)
{
    switch (kind) {
    case WF_WAVE_GAUSSIAN: {
        const float arg = (t - p0) / p1;
        return std::exp(-0.5f * arg * arg);
    }
    case WF_WAVE_SINUSOID:
        return std::sin(2.0f * kPiF * p0 * t);
    case WF_WAVE_RICKER: {
        const float arg = kPiF * p0 * (t - p1);
        const float sq = arg * arg;
        return (1.0f - 2.0f * sq) * std::exp(-sq);
    }
    default:
        return 0.0f;
    }
}

// src/source.rs:79-95 (value only) at t = (step+1)*dt, src/engine.rs:769
int wf_source_values(uint32_t n, const wf_source_desc *src, float dt, uint64_t first_step,
                     uint64_t nsteps, float *values)
{
    if (n && (!src || !values)) return fail(WF_ERR_INVALID, "wf_source_values: null pointer");
    for (uint64_t k = 0; k < nsteps; ++k) {
        const float t = (static_cast<float>(first_step + k) + 1.0f) * dt;
        for (uint32_t s = 0; s < n; ++s) {
            const wf_source_desc &d = src[s];
            values[k * n + s] = d.amplitude * wf_waveform(d.waveform, d.p0, d.p1, t);
        }
    }
    return WF_OK;
}

// src/probe.rs:126-138 (twiddles and the start gate)
int wf_probe_twiddles(uint32_t n, const wf_probe_site *sites, const float *freqs,
                      const float *starts, float dt, uint64_t first_step, uint64_t nsteps,
                      double *cos_tab, double *sin_tab, uint8_t *active)
{
    if (n && (!sites || !starts || !active))
        return fail(WF_ERR_INVALID, "wf_probe_twiddles: null pointer");
    uint64_t nf = 0;
    for (uint32_t p = 0; p < n; ++p) nf += sites[p].nfreq;
    if (nf && (!freqs || !cos_tab || !sin_tab))
        return fail(WF_ERR_INVALID, "wf_probe_twiddles: null frequency table");
    std::vector<double> omega(nf);
    for (uint64_t j = 0; j < nf; ++j) omega[j] = 2.0 * M_PI * static_cast<double>(freqs[j]);
    for (uint64_t k = 0; k < nsteps; ++k) {
        const float t = (static_cast<float>(first_step + k) + 1.0f) * dt;
        for (uint32_t p = 0; p < n; ++p) active[k * n + p] = (t < starts[p]) ? 0 : 1;
        for (uint64_t j = 0; j < nf; ++j) {
            const double phase = omega[j] * static_cast<double>(t);
            cos_tab[k * nf + j] = std::cos(phase);
            sin_tab[k * nf + j] = std::sin(phase);
        }
    }
    return WF_OK;
}

// src/probe.rs:143-167
void wf_probe_spectrum(uint32_t nfreq, const double *sum_cos, const double *sum_sin,
                       uint64_t n, double *amplitude, double *phase_rad)
{
    const double nn = static_cast<double>(n > 0 ? n : 1);
    for (uint32_t i = 0; i < nfreq; ++i) {
        const double a = 2.0 * sum_cos[i] / nn;
        const double b = 2.0 * sum_sin[i] / nn;
        amplitude[i] = std::sqrt(a * a + b * b);
        phase_rad[i] = std::atan2(a, b);
    }
}

// src/main.rs:403-425
void wf_voxelize_demo_sphere(uint32_t nx, uint32_t ny, uint32_t nz, uint8_t *ids)
{
    const float cx = static_cast<float>(nx) / 2.0f, cy = static_cast<float>(ny) / 2.0f,
                cz = static_cast<float>(nz) / 2.0f;
    uint32_t mn = nx < ny ? nx : ny;
    mn = mn < nz ? mn : nz;
    const float radius = static_cast<float>(mn) / 4.0f;
    const float r2 = radius * radius;
    for (uint32_t z = 0; z < nz; ++z)
        for (uint32_t y = 0; y < ny; ++y)
            for (uint32_t x = 0; x < nx; ++x) {
                const float dx = static_cast<float>(x) + 0.5f - cx;
                const float dy = static_cast<float>(y) + 0.5f - cy;
                const float dz = static_cast<float>(z) + 0.5f - cz;
                if (dx * dx + dy * dy + dz * dz <= r2)
                    ids[(static_cast<size_t>(z) * ny + y) * nx + x] = 1;
            }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Scene files: src/scene.rs:81-230
// ---------------------------------------------------------------------------
namespace {

struct Primitive {
    bool is_box;
    float v[9];  // eps_r, mu_r, sigma, then centre+radius or lo+hi
};

// Voxels [lo, hi) along one axis whose own term can pass the sphere
// predicate, fl(d * d) <= r2 with d = fl(fl(v + 0.5) - c).  The predicate
// (src/scene.rs:197-209) compares a rounded sum of three non-negative rounded
// squares with r2, and such a sum is never below any of its terms, so a voxel
// whose term on one axis exceeds r2 fails whatever the other two are.  d is
// monotone in v, so the passing voxels form one interval, found here by
// testing every v of the axis: O(n) per axis instead of O(n^3), and the box
// [lo, hi)^3 holds every voxel the reference's full scan sets, for any
// centre and radius (including spheres far larger than the grid, whose f32
// rounding at |c| spans many cells).
void sphere_range(float c, float r2, uint32_t n, uint32_t *lo, uint32_t *hi)
{
    uint32_t a = n, b = n;
    for (uint32_t v = 0; v < n; ++v) {
        const float d = static_cast<float>(v) + 0.5f - c;
        if (d * d <= r2) {
            if (a == n) a = v;
            b = v + 1;
        }
    }
    *lo = a == n ? 0 : a;
    *hi = a == n ? 0 : b;
}

// The token grammar of Rust's f32 FromStr, which the reference parses scene
// numbers with (src/scene.rs:96): [+-]? then inf | infinity | nan | a decimal
// (D+ | D+ '.' D* | D* '.' D+, exponent [eE][+-]?D+ optional), letters in any
// case.  strtof reads more (hex floats, nan(...)); those tokens are errors.
bool decimal_token(const char *s)
{
    if (*s == '+' || *s == '-') ++s;
    if (strcasecmp(s, "inf") == 0 || strcasecmp(s, "infinity") == 0 || strcasecmp(s, "nan") == 0)
        return true;
    int digits = 0;
    while (std::isdigit(static_cast<unsigned char>(*s))) ++s, ++digits;
    if (*s == '.') {
        ++s;
        while (std::isdigit(static_cast<unsigned char>(*s))) ++s, ++digits;
    }
    if (digits == 0) return false;
    if (*s == 'e' || *s == 'E') {
        ++s;
        if (*s == '+' || *s == '-') ++s;
        if (!std::isdigit(static_cast<unsigned char>(*s))) return false;
        while (std::isdigit(static_cast<unsigned char>(*s))) ++s;
    }
    return *s == 0;
}

int parse_scene(const char *text, std::vector<Primitive> &out)
{
    size_t line_no = 0;
    const char *p = text;
    while (*p) {
        const char *e = std::strchr(p, '\n');
        std::string line(p, e ? static_cast<size_t>(e - p) : std::strlen(p));
        ++line_no;
        std::vector<std::string> tok;
        size_t i = 0;
        while (i < line.size()) {
            while (i < line.size() && std::isspace(static_cast<unsigned char>(line[i]))) ++i;
            size_t j = i;
            while (j < line.size() && !std::isspace(static_cast<unsigned char>(line[j]))) ++j;
            if (j > i) tok.emplace_back(line.substr(i, j - i));
            i = j;
        }
        if (!tok.empty() && tok[0][0] != '#') {
            const bool sphere = tok[0] == "sphere", box = tok[0] == "box";
            if (!sphere && !box)
                return fail(WF_ERR_INVALID, "scene line %zu: unknown primitive '%s'", line_no,
                            tok[0].c_str());
            const size_t want = sphere ? 7 : 9;
            if (tok.size() - 1 != want)
                return fail(WF_ERR_INVALID, "scene line %zu: expected %zu values, got %zu",
                            line_no, want, tok.size() - 1);
            Primitive pr{};
            pr.is_box = box;
            for (size_t k = 0; k < want; ++k) {
                const char *s = tok[k + 1].c_str();
                char *end = nullptr;
                const float f = std::strtof(s, &end);
                if (!decimal_token(s) || end == s || *end != 0)
                    return fail(WF_ERR_INVALID, "scene line %zu: invalid number '%s'",
                                line_no, s);
                pr.v[k] = f;
            }
            out.push_back(pr);
        }
        if (!e) break;
        p = e + 1;
    }
    return WF_OK;
}

}  // namespace

extern "C" int wf_scene_voxelize(const char *text, uint32_t nx, uint32_t ny, uint32_t nz,
                                 float dt, float dx, uint8_t *ids, int32_t *table)
{
    if (!text || !ids || !table) return fail(WF_ERR_INVALID, "wf_scene_voxelize: null pointer");
    std::vector<Primitive> prims;
    if (int rc = parse_scene(text, prims)) return rc;

    // Deduplicate (eps_r, mu_r, sigma) by exact bit pattern, ids 1.. in
    // first-seen order (src/scene.rs:30-49, 157-183).
    struct KeyHash {
        size_t operator()(const std::string &k) const { return std::hash<std::string>()(k); }
    };
    std::unordered_map<std::string, uint8_t, KeyHash> ids_by_key;
    unsigned next_id = 1;
    for (const Primitive &pr : prims) {
        const std::string key(reinterpret_cast<const char *>(pr.v), 3 * sizeof(float));
        uint8_t id;
        auto it = ids_by_key.find(key);
        if (it != ids_by_key.end()) {
            id = it->second;
        } else {
            if (next_id > 255) return fail(WF_ERR_INVALID, "scene uses more than 255 distinct materials");
            id = static_cast<uint8_t>(next_id++);
            wf_material_coeffs(pr.v[0], pr.v[1], pr.v[2], dt, dx, table + 4 * id);
            ids_by_key.emplace(key, id);
        }
        if (!pr.is_box) {
            const float cx = pr.v[3], cy = pr.v[4], cz = pr.v[5], r = pr.v[6];
            const float r2 = r * r;
            uint32_t x0, x1, y0, y1, z0, z1;
            sphere_range(cx, r2, nx, &x0, &x1);
            sphere_range(cy, r2, ny, &y0, &y1);
            sphere_range(cz, r2, nz, &z0, &z1);
            for (uint32_t z = z0; z < z1; ++z)
                for (uint32_t y = y0; y < y1; ++y) {
                    uint8_t *row = ids + (static_cast<size_t>(z) * ny + y) * nx;
                    for (uint32_t x = x0; x < x1; ++x) {
                        const float ddx = static_cast<float>(x) + 0.5f - cx;
                        const float ddy = static_cast<float>(y) + 0.5f - cy;
                        const float ddz = static_cast<float>(z) + 0.5f - cz;
                        if (ddx * ddx + ddy * ddy + ddz * ddz <= r2) row[x] = id;
                    }
                }
        } else {
            const uint32_t n[3] = {nx, ny, nz};
            size_t lo[3], hi[3];
            for (int ax = 0; ax < 3; ++ax) {
                const float a = pr.v[3 + ax], b = pr.v[6 + ax];
                lo[ax] = sat_usize(std::floor(std::fmax(std::fmin(a, b), 0.0f)));
                const size_t h = sat_usize(std::fmax(std::ceil(std::fmax(a, b)), 0.0f));
                hi[ax] = h < n[ax] ? h : n[ax];
            }
            for (size_t z = lo[2]; z < hi[2]; ++z)
                for (size_t y = lo[1]; y < hi[1]; ++y) {
                    uint8_t *row = ids + (z * ny + y) * nx;
                    for (size_t x = lo[0]; x < hi[0]; ++x) row[x] = id;
                }
        }
    }
    return static_cast<int>(ids_by_key.size());
}

// `wavefront` command-line driver on the MI355X engine.
//
// Same flags, defaults, messages and exit codes as the reference CLI
// (src/main.rs:131-619): the demo sphere or a --scene file, sources and
// probes from the shorthand or key=value flags, the mmap'd material grid
// file, engine::run with snapshot egress (here wf_run on the GPU), and the
// probe spectra printed at the end.  Host-side tables (materials, CPML
// profiles, waveform values, probe twiddles, spectra) come from the library's
// host builders, so every printed number follows the reference arithmetic.
#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "wavefront_hip.h"

namespace {

// ---------------------------------------------------------------------------
// Rust formatting used by the reference's messages
// ---------------------------------------------------------------------------

// Shortest decimal digits that read back as exactly v (f32): returns the
// digit string and the decimal exponent of its first digit.
void shortest_f32(float v, std::string &digits, int &exp10)
{
    const double x = std::fabs((double)v);
    for (int p = 0; p <= 9; ++p) {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%.*e", p, x);
        // mantissa digits and exponent of the correctly rounded p-digit form
        std::string d;
        const char *s = buf;
        for (; *s && *s != 'e'; ++s)
            if (*s != '.') d += *s;
        const int e = std::atoi(s + 1);
        // the correctly rounded form, then its two neighbours (the closest
        // p-digit decimal can fall outside the narrower half of the
        // round-trip interval just above a power of two)
        const long long m = std::atoll(d.c_str());
        long long best = -1;
        double best_err = 0.0;
        for (long long cand : {m, m - 1, m + 1}) {
            if (cand <= 0) continue;
            char t[64];
            std::snprintf(t, sizeof t, "%llde%d", cand, e - p);
            if (std::strtof(t, nullptr) != (float)x) continue;
            const double err = std::fabs(std::strtod(t, nullptr) - x);
            if (best < 0 || err < best_err) {
                best = cand;
                best_err = err;
            }
        }
        if (best < 0) continue;
        digits = std::to_string(best);
        exp10 = e - p + (int)digits.size() - 1;
        while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
        return;
    }
    digits = "0";
    exp10 = 0;
}

// Rust's `{:?}` for f32: shortest round-trip digits, positional with at
// least one fractional digit when 1e-4 <= |v| < 1e16 (or v == 0), otherwise
// d.ddde<exp>.
std::string rust_debug_f32(float v)
{
    if (std::isnan(v)) return "NaN";
    const std::string sign = std::signbit(v) ? "-" : "";
    if (std::isinf(v)) return sign + "inf";
    const float a = std::fabs(v);
    if (a == 0.0f) return sign + "0.0";
    std::string d;
    int e = 0;
    shortest_f32(a, d, e);
    if (a >= 1e-4f && a < 1e16f) {
        std::string out;
        if (e >= 0) {
            std::string ip = d.substr(0, std::min<size_t>(d.size(), (size_t)e + 1));
            while ((int)ip.size() < e + 1) ip += '0';
            std::string fp = (size_t)e + 1 < d.size() ? d.substr(e + 1) : "0";
            out = ip + "." + fp;
        } else {
            out = "0." + std::string((size_t)(-e - 1), '0') + d;
        }
        return sign + out;
    }
    std::string out = d.substr(0, 1);
    if (d.size() > 1) out += "." + d.substr(1);
    return sign + out + "e" + std::to_string(e);
}

// Rust's `{:.Ne}`: C's %.Ne with the exponent as a plain signed integer.
std::string rust_exp(double v, int prec)
{
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v < 0 ? "-inf" : "inf";
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.*e", prec, v);
    char *e = std::strchr(buf, 'e');
    const int ex = std::atoi(e + 1);
    *e = '\0';
    return std::string(buf) + "e" + std::to_string(ex);
}

std::string rust_fixed(double v, int prec)
{
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v < 0 ? "-inf" : "inf";
    char buf[512];
    std::snprintf(buf, sizeof buf, "%.*f", prec, v);
    return buf;
}

// Rust's `{:?}` for a path: a quoted, escaped string.
std::string rust_debug_path(const std::string &p)
{
    std::string out = "\"";
    for (unsigned char ch : p) {
        switch (ch) {
        case '"': out += "\\\""; break;
        case '\\': out += "\\\\"; break;
        case '\n': out += "\\n"; break;
        case '\r': out += "\\r"; break;
        case '\t': out += "\\t"; break;
        default:
            if (ch < 0x20 || ch == 0x7f) {
                char b[16];
                std::snprintf(b, sizeof b, "\\u{%x}", ch);
                out += b;
            } else {
                out += (char)ch;
            }
        }
    }
    return out + "\"";
}

// io::Error's Display for an OS error.
std::string os_error(int err)
{
    return std::string(std::strerror(err)) + " (os error " + std::to_string(err) + ")";
}

const char *component_name(int c)
{
    static const char *n[6] = {"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"};
    return (c >= 0 && c < 6) ? n[c] : "?";
}

// ---------------------------------------------------------------------------
// Argument parsing (src/main.rs:38-398)
// ---------------------------------------------------------------------------

struct SourceSpec {
    bool has_x = false, has_y = false, has_z = false;
    uint64_t x = 0, y = 0, z = 0;
    int component = WF_EZ;
    int waveform = WF_WAVE_RICKER;
    bool has_freq = false;
    float freq = 0.0f;
    float amplitude = 1.0f;
};

struct ProbeSpec {
    bool has_x = false, has_y = false, has_z = false;
    uint64_t x = 0, y = 0, z = 0;
    int component = WF_EZ;
    std::vector<float> freqs;
    float start = 0.0f;
};

struct Config {
    uint64_t nx = 64, ny = 64, nz = 64;
    float dx = 1.0e-3f;
    float dt = 1.5e-12f;
    uint64_t num_steps = 200;
    uint64_t snapshot_every = 20;
    uint64_t pml_thickness = WF_BLOCK_DIM;  // PmlConfig::default().thickness
    bool has_scene = false;
    std::string scene_path;
    std::vector<SourceSpec> sources;
    std::vector<ProbeSpec> probes;
    std::string materials_path = "materials.grid";
    std::string output_path = "wave_trajectory.bin";
};

struct ParseError {
    std::string msg;
};

// usize::from_str: optional '+', decimal digits, no overflow.
bool parse_usize(const std::string &s, uint64_t &out)
{
    size_t i = 0;
    if (i < s.size() && s[i] == '+') ++i;
    if (i == s.size()) return false;
    uint64_t v = 0;
    for (; i < s.size(); ++i) {
        if (s[i] < '0' || s[i] > '9') return false;
        const uint64_t d = (uint64_t)(s[i] - '0');
        if (v > (UINT64_MAX - d) / 10) return false;
        v = v * 10 + d;
    }
    out = v;
    return true;
}

// f32::from_str: [+-] then inf | infinity | nan (any case) or
// digits[.digits][(e|E)[+-]digits] with at least one mantissa digit;
// correctly rounded (glibc strtof is).
bool parse_f32(const std::string &s, float &out)
{
    size_t i = 0;
    if (i < s.size() && (s[i] == '+' || s[i] == '-')) ++i;
    std::string rest = s.substr(i);
    std::string low;
    for (char ch : rest) low += (char)std::tolower((unsigned char)ch);
    if (low == "inf" || low == "infinity" || low == "nan") {
        out = std::strtof(s.c_str(), nullptr);
        return true;
    }
    size_t j = i, mant = 0;
    while (j < s.size() && std::isdigit((unsigned char)s[j])) ++j, ++mant;
    if (j < s.size() && s[j] == '.') {
        ++j;
        while (j < s.size() && std::isdigit((unsigned char)s[j])) ++j, ++mant;
    }
    if (mant == 0) return false;
    if (j < s.size() && (s[j] == 'e' || s[j] == 'E')) {
        ++j;
        if (j < s.size() && (s[j] == '+' || s[j] == '-')) ++j;
        size_t ed = 0;
        while (j < s.size() && std::isdigit((unsigned char)s[j])) ++j, ++ed;
        if (ed == 0) return false;
    }
    if (j != s.size()) return false;
    out = std::strtof(s.c_str(), nullptr);
    return true;
}

std::string trim(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

std::vector<std::string> split(const std::string &s, char sep)
{
    std::vector<std::string> out;
    size_t a = 0;
    for (;;) {
        const size_t b = s.find(sep, a);
        out.push_back(s.substr(a, b == std::string::npos ? std::string::npos : b - a));
        if (b == std::string::npos) break;
        a = b + 1;
    }
    return out;
}

int parse_source_component(const std::string &v)
{
    if (v == "ex") return WF_EX;
    if (v == "ey") return WF_EY;
    if (v == "ez") return WF_EZ;
    throw ParseError{"invalid source component: " + v + " (expected ex, ey, or ez)"};
}

int parse_probe_component(const std::string &v)
{
    static const char *n[6] = {"ex", "ey", "ez", "hx", "hy", "hz"};
    for (int c = 0; c < 6; ++c)
        if (v == n[c]) return c;
    throw ParseError{"invalid probe component: " + v + " (expected ex, ey, ez, hx, hy, or hz)"};
}

int parse_waveform_kind(const std::string &v)
{
    if (v == "gaussian") return WF_WAVE_GAUSSIAN;
    if (v == "sinusoid") return WF_WAVE_SINUSOID;
    if (v == "ricker") return WF_WAVE_RICKER;
    throw ParseError{"invalid waveform: " + v + " (expected gaussian, sinusoid, or ricker)"};
}

// src/main.rs:182-211
SourceSpec parse_source_spec(const std::string &s)
{
    SourceSpec spec;
    for (const std::string &raw : split(s, ',')) {
        const std::string pair = trim(raw);
        if (pair.empty()) continue;
        const size_t eq = pair.find('=');
        if (eq == std::string::npos) throw ParseError{"--source: expected key=value, got '" + pair + "'"};
        const std::string key = trim(pair.substr(0, eq)), value = trim(pair.substr(eq + 1));
        auto num = [&](uint64_t &dst, bool &has, const char *name) {
            if (!parse_usize(value, dst))
                throw ParseError{std::string("--source: invalid ") + name + " '" + value + "'"};
            has = true;
        };
        if (key == "x") num(spec.x, spec.has_x, "x");
        else if (key == "y") num(spec.y, spec.has_y, "y");
        else if (key == "z") num(spec.z, spec.has_z, "z");
        else if (key == "component") spec.component = parse_source_component(value);
        else if (key == "waveform") spec.waveform = parse_waveform_kind(value);
        else if (key == "freq") {
            if (!parse_f32(value, spec.freq)) throw ParseError{"--source: invalid freq '" + value + "'"};
            spec.has_freq = true;
        } else if (key == "amplitude") {
            if (!parse_f32(value, spec.amplitude))
                throw ParseError{"--source: invalid amplitude '" + value + "'"};
        } else {
            throw ParseError{"--source: unknown key '" + key + "'"};
        }
    }
    return spec;
}

// src/main.rs:217-243
ProbeSpec parse_probe_spec(const std::string &s)
{
    ProbeSpec spec;
    for (const std::string &raw : split(s, ',')) {
        const std::string pair = trim(raw);
        if (pair.empty()) continue;
        const size_t eq = pair.find('=');
        if (eq == std::string::npos) throw ParseError{"--probe: expected key=value, got '" + pair + "'"};
        const std::string key = trim(pair.substr(0, eq)), value = trim(pair.substr(eq + 1));
        auto num = [&](uint64_t &dst, bool &has, const char *name) {
            if (!parse_usize(value, dst))
                throw ParseError{std::string("--probe: invalid ") + name + " '" + value + "'"};
            has = true;
        };
        if (key == "x") num(spec.x, spec.has_x, "x");
        else if (key == "y") num(spec.y, spec.has_y, "y");
        else if (key == "z") num(spec.z, spec.has_z, "z");
        else if (key == "component") spec.component = parse_probe_component(value);
        else if (key == "freq") {
            std::vector<float> fs;
            for (const std::string &f : split(value, ';')) {
                float v;
                if (!parse_f32(trim(f), v)) throw ParseError{"--probe: invalid freq '" + f + "'"};
                fs.push_back(v);
            }
            spec.freqs = fs;
        } else if (key == "start") {
            if (!parse_f32(value, spec.start)) throw ParseError{"--probe: invalid start '" + value + "'"};
        } else {
            throw ParseError{"--probe: unknown key '" + key + "'"};
        }
    }
    return spec;
}

void print_usage()
{
    std::fprintf(stderr, "%s",
        "wavefront -- out-of-core 3D FDTD electromagnetic simulator\n\n"
        "USAGE:\n    wavefront [OPTIONS]\n\n"
        "OPTIONS:\n"
        "    --nx <N>               grid size along X in voxels (multiple of 8) [default: 64]\n"
        "    --ny <N>               grid size along Y in voxels (multiple of 8) [default: 64]\n"
        "    --nz <N>               grid size along Z in voxels (multiple of 8) [default: 64]\n"
        "    --dx <METERS>          uniform cell size [default: 1.0e-3]\n"
        "    --dt <SECONDS>         timestep [default: 1.5e-12]\n"
        "    --steps <N>            number of timesteps to run [default: 200]\n"
        "    --snapshot-every <N>   timesteps between snapshot writes [default: 20]\n"
        "    --pml-thickness <N>    absorbing boundary depth in voxels, 0 disables it [default: 8]\n"
        "    --scene <PATH>         plain-text scene file (see src/scene.rs); omit for the demo sphere\n"
        "    --source-x/-y/-z <N>   first source's voxel position [default: domain center]\n"
        "    --source-component <C> first source's field component: ex, ey, ez [default: ez]\n"
        "    --source-waveform <W>  first source's waveform: gaussian, sinusoid, or ricker [default: ricker]\n"
        "    --source-freq <HZ>     first source's drive frequency [default: 1/(20*dt)]\n"
        "    --source-amplitude <A> first source's peak amplitude [default: 1.0]\n"
        "    --source <SPEC>        add another source: key=value,... (x,y,z,component,waveform,freq,amplitude)\n"
        "    --probe-x/-y/-z <N>    first probe's voxel position; enables it together with --probe-freq\n"
        "    --probe-component <C>  first probe's field component: ex, ey, ez, hx, hy, hz [default: ez]\n"
        "    --probe-freq <HZ,...>  comma-separated frequencies (Hz) the first probe's running DFT tracks\n"
        "    --probe-start <SECONDS> first probe's ignore-samples-before time [default: 0.0]\n"
        "    --probe <SPEC>         add another probe: key=value,... (x,y,z,component,freq[;freq...],start)\n"
        "    --materials <PATH>     backing file for the mmap'd material grid [default: materials.grid]\n"
        "    --output <PATH>        Direct I/O snapshot stream path [default: wave_trajectory.bin]\n"
        "    -h, --help             print this message\n\n"
        "Repeat --source/--probe to configure more than one; the individual\n"
        "--source-*/--probe-* flags configure only the first. --probe's freq\n"
        "list uses ';' (quote the whole value in your shell, e.g.\n"
        "--probe \"x=10,y=10,z=10,freq=3e10;6e10\").\n");
}

Config parse_args(int argc, char **argv)
{
    Config c;
    auto first_source = [&]() -> SourceSpec & {
        if (c.sources.empty()) c.sources.emplace_back();
        return c.sources[0];
    };
    auto first_probe = [&]() -> ProbeSpec & {
        if (c.probes.empty()) c.probes.emplace_back();
        return c.probes[0];
    };
    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        auto next_value = [&](const std::string &name) -> std::string {
            if (i + 1 >= argc) throw ParseError{"missing value for " + name};
            return argv[++i];
        };
        auto num = [&](const std::string &name) -> uint64_t {
            const std::string s = next_value(name);
            uint64_t v;
            if (!parse_usize(s, v)) throw ParseError{"invalid integer for " + name + ": " + s};
            return v;
        };
        auto flt = [&](const std::string &name, const std::string &s) -> float {
            float v;
            if (!parse_f32(s, v)) throw ParseError{"invalid number for " + name + ": " + s};
            return v;
        };
        if (arg == "--nx") c.nx = num(arg);
        else if (arg == "--ny") c.ny = num(arg);
        else if (arg == "--nz") c.nz = num(arg);
        else if (arg == "--dx") c.dx = flt(arg, next_value(arg));
        else if (arg == "--dt") c.dt = flt(arg, next_value(arg));
        else if (arg == "--steps") c.num_steps = num(arg);
        else if (arg == "--snapshot-every") c.snapshot_every = num(arg);
        else if (arg == "--pml-thickness") c.pml_thickness = num(arg);
        else if (arg == "--scene") {
            c.scene_path = next_value(arg);
            c.has_scene = true;
        } else if (arg == "--source-x") {
            const uint64_t v = num(arg);
            first_source().x = v, first_source().has_x = true;
        } else if (arg == "--source-y") {
            const uint64_t v = num(arg);
            first_source().y = v, first_source().has_y = true;
        } else if (arg == "--source-z") {
            const uint64_t v = num(arg);
            first_source().z = v, first_source().has_z = true;
        } else if (arg == "--source-component") {
            const int v = parse_source_component(next_value(arg));
            first_source().component = v;
        } else if (arg == "--source-waveform") {
            const int v = parse_waveform_kind(next_value(arg));
            first_source().waveform = v;
        } else if (arg == "--source-freq") {
            const float v = flt(arg, next_value(arg));
            first_source().freq = v, first_source().has_freq = true;
        } else if (arg == "--source-amplitude") {
            const float v = flt(arg, next_value(arg));
            first_source().amplitude = v;
        } else if (arg == "--source") {
            c.sources.push_back(parse_source_spec(next_value(arg)));
        } else if (arg == "--probe-x") {
            const uint64_t v = num(arg);
            first_probe().x = v, first_probe().has_x = true;
        } else if (arg == "--probe-y") {
            const uint64_t v = num(arg);
            first_probe().y = v, first_probe().has_y = true;
        } else if (arg == "--probe-z") {
            const uint64_t v = num(arg);
            first_probe().z = v, first_probe().has_z = true;
        } else if (arg == "--probe-component") {
            const int v = parse_probe_component(next_value(arg));
            first_probe().component = v;
        } else if (arg == "--probe-freq") {
            std::vector<float> fs;
            for (const std::string &s : split(next_value(arg), ',')) fs.push_back(flt(arg, s));
            first_probe().freqs = fs;
        } else if (arg == "--probe-start") {
            const float v = flt(arg, next_value(arg));
            first_probe().start = v;
        } else if (arg == "--probe") {
            c.probes.push_back(parse_probe_spec(next_value(arg)));
        } else if (arg == "--materials") {
            c.materials_path = next_value(arg);
        } else if (arg == "--output") {
            c.output_path = next_value(arg);
        } else if (arg == "-h" || arg == "--help") {
            print_usage();
            std::exit(0);
        } else {
            throw ParseError{"unrecognized argument: " + arg};
        }
    }
    return c;
}

// ---------------------------------------------------------------------------
// Setup and run (src/main.rs:401-600)
// ---------------------------------------------------------------------------

struct RunError {
    std::string msg;
    int exit_code = 1;
};

// build_one_source, src/main.rs:431-460
wf_source_desc build_one_source(const SourceSpec &s, float dt, const Config &c)
{
    wf_source_desc d{};
    d.site.x = (uint32_t)(s.has_x ? s.x : c.nx / 2);
    d.site.y = (uint32_t)(s.has_y ? s.y : c.ny / 2);
    d.site.z = (uint32_t)(s.has_z ? s.z : c.nz / 2);
    d.site.component = s.component;
    d.site.hard = 0;
    d.amplitude = s.amplitude;
    d.waveform = s.waveform;
    const float freq = s.has_freq ? s.freq : 1.0f / (20.0f * dt);
    if (s.waveform == WF_WAVE_SINUSOID) {
        d.p0 = freq;
    } else if (s.waveform == WF_WAVE_RICKER) {
        d.p0 = freq;
        d.p1 = 1.0f / freq;
    } else {
        const float spread = 0.5f / freq;
        d.p0 = 4.0f * spread;
        d.p1 = spread;
    }
    return d;
}

std::string waveform_debug(const wf_source_desc &d)
{
    if (d.waveform == WF_WAVE_SINUSOID) return "Sinusoid { freq_hz: " + rust_debug_f32(d.p0) + " }";
    if (d.waveform == WF_WAVE_RICKER)
        return "RickerWavelet { peak_freq_hz: " + rust_debug_f32(d.p0) + ", t0: " + rust_debug_f32(d.p1) + " }";
    return "GaussianPulse { t0: " + rust_debug_f32(d.p0) + ", spread: " + rust_debug_f32(d.p1) + " }";
}

void check(int rc, const char *what)
{
    if (rc < 0) throw RunError{std::string(what) + wf_last_error()};
}

bool read_file(const std::string &path, std::string &out, int &err)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) {
        err = errno;
        return false;
    }
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, n);
    const bool ok = !std::ferror(f);
    err = ok ? 0 : EIO;
    std::fclose(f);
    return ok;
}

void run(const Config &c)
{
    if (c.nx % WF_BLOCK_DIM || c.ny % WF_BLOCK_DIM || c.nz % WF_BLOCK_DIM || !c.nx || !c.ny || !c.nz ||
        c.nx > UINT32_MAX || c.ny > UINT32_MAX || c.nz > UINT32_MAX)
        // GridDims::new's assertion (src/layout.rs:60-66); a failed assert
        // exits with Rust's panic status
        throw RunError{"grid dimensions must be multiples of BLOCK_DIM (8) so the field grid tiles "
                       "exactly with no partial edge blocks",
                       101};
    const uint64_t voxels = c.nx * c.ny * c.nz;
    const uint64_t field_mib = voxels * 24 / (1024 * 1024);
    std::fprintf(stderr,
                 "wavefront: %llux%llux%llu voxels (%llu MiB field state, dx=%s m, dt=%s s, %llu steps, "
                 "pml=%llu voxels)\n",
                 (unsigned long long)c.nx, (unsigned long long)c.ny, (unsigned long long)c.nz,
                 (unsigned long long)field_mib, rust_exp(c.dx, 3).c_str(), rust_exp(c.dt, 3).c_str(),
                 (unsigned long long)c.num_steps, (unsigned long long)c.pml_thickness);

    // MaterialGrid::create (src/layout.rs:291-308): the backing file exists
    // from here on, sized to one byte per voxel
    const int mfd = ::open(c.materials_path.c_str(), O_RDWR | O_CREAT | O_TRUNC, 0644);
    if (mfd < 0 || ::ftruncate(mfd, (off_t)voxels) != 0) {
        const int err = errno;
        if (mfd >= 0) ::close(mfd);
        throw RunError{"failed to create material grid at " + rust_debug_path(c.materials_path) + ": " +
                       os_error(err)};
    }
    std::vector<uint8_t> ids(voxels, 0);
    std::vector<int32_t> table(256 * 4);
    wf_material_table_vacuum(c.dt, c.dx, table.data());
    if (c.has_scene) {
        std::string text;
        int err = 0;
        if (!read_file(c.scene_path, text, err)) {
            ::close(mfd);
            throw RunError{"failed to read scene file " + rust_debug_path(c.scene_path) + ": " + os_error(err)};
        }
        const int n = wf_scene_voxelize(text.c_str(), (uint32_t)c.nx, (uint32_t)c.ny, (uint32_t)c.nz, c.dt,
                                        c.dx, ids.data(), table.data());
        if (n < 0) {
            ::close(mfd);
            throw RunError{wf_last_error()};
        }
        std::fprintf(stderr, "wavefront: loaded scene %s (%d distinct materials)\n",
                     rust_debug_path(c.scene_path).c_str(), n);
    } else {
        wf_voxelize_demo_sphere((uint32_t)c.nx, (uint32_t)c.ny, (uint32_t)c.nz, ids.data());
        wf_material_coeffs(4.0f, 1.0f, 0.0f, c.dt, c.dx, &table[1 * 4]);
    }
    // flush
    {
        uint64_t off = 0;
        while (off < voxels) {
            const size_t len = (size_t)std::min<uint64_t>(voxels - off, 1ull << 30);
            const ssize_t w = ::pwrite(mfd, ids.data() + off, len, (off_t)off);
            if (w <= 0) {
                const int err = w < 0 ? errno : EIO;
                ::close(mfd);
                throw RunError{"failed to flush material grid: " + os_error(err)};
            }
            off += (uint64_t)w;
        }
        if (::close(mfd) != 0) throw RunError{"failed to flush material grid: " + os_error(errno)};
    }

    std::vector<wf_source_desc> sources;
    if (c.sources.empty())
        sources.push_back(build_one_source(SourceSpec{}, c.dt, c));
    else
        for (const SourceSpec &s : c.sources) sources.push_back(build_one_source(s, c.dt, c));
    for (const wf_source_desc &d : sources)
        std::fprintf(stderr, "wavefront: source at (%u, %u, %u) driving %s, %s\n", d.site.x, d.site.y,
                     d.site.z, component_name(d.site.component), waveform_debug(d).c_str());

    // build_one_probe, src/main.rs:479-500
    std::vector<wf_probe_site> psites;
    std::vector<float> freqs, starts;
    for (const ProbeSpec &p : c.probes) {
        if (!p.has_x || !p.has_y || !p.has_z)
            throw RunError{"each probe requires x, y, and z (via --probe-x/-y/-z, or x=/y=/z= in --probe)"};
        if (p.freqs.empty())
            throw RunError{"each probe requires at least one frequency (via --probe-freq, or freq= in --probe)"};
        if (p.x >= c.nx || p.y >= c.ny || p.z >= c.nz)
            throw RunError{"probe position (" + std::to_string(p.x) + ", " + std::to_string(p.y) + ", " +
                           std::to_string(p.z) + ") is outside the " + std::to_string(c.nx) + "x" +
                           std::to_string(c.ny) + "x" + std::to_string(c.nz) + " domain"};
        psites.push_back(wf_probe_site{(uint32_t)p.x, (uint32_t)p.y, (uint32_t)p.z, p.component,
                                       (uint32_t)p.freqs.size()});
        freqs.insert(freqs.end(), p.freqs.begin(), p.freqs.end());
        starts.push_back(p.start);
    }
    for (const wf_probe_site &p : psites)
        std::fprintf(stderr, "wavefront: probe at (%u, %u, %u) tracking %s\n", p.x, p.y, p.z,
                     component_name(p.component));

    // PmlContext::build with PmlConfig { thickness, ..default } (src/main.rs:563-568)
    uint32_t shell[3] = {0, 0, 0};
    std::vector<float> prof;
    const bool pml = c.pml_thickness > 0;
    if (pml) {
        const uint32_t tb = wf_effective_pml_thickness_blocks(
            (uint32_t)c.nx, (uint32_t)c.ny, (uint32_t)c.nz,
            (uint32_t)std::min<uint64_t>(c.pml_thickness, UINT32_MAX));
        shell[0] = shell[1] = shell[2] = tb;
        const uint64_t n[3] = {c.nx, c.ny, c.nz};
        for (int a = 0; a < 3; ++a) {
            std::vector<float> b(n[a]), al(n[a]), ik(n[a]);
            wf_pml_profile((uint32_t)n[a], 3.0f, 7.0f, 0.05f, 1.0e-6f, c.dt, c.dx, tb * WF_BLOCK_DIM, b.data(),
                           al.data(), ik.data());
            prof.insert(prof.end(), b.begin(), b.end());
            prof.insert(prof.end(), al.begin(), al.end());
            prof.insert(prof.end(), ik.begin(), ik.end());
        }
    }
    const uint64_t every = c.snapshot_every ? c.snapshot_every : 1;

    // engine::run
    const char *pre = "simulation run failed: ";
    const uint64_t ns = c.num_steps, nsrc = sources.size(), nprobe = psites.size();
    std::vector<float> values(ns * nsrc);
    check(wf_source_values((uint32_t)nsrc, sources.data(), c.dt, 0, ns, values.data()), pre);
    const uint64_t nf = freqs.size();
    std::vector<double> ctab(ns * nf), stab(ns * nf);
    std::vector<uint8_t> active(ns * nprobe);
    if (nprobe)
        check(wf_probe_twiddles((uint32_t)nprobe, psites.data(), freqs.data(), starts.data(), c.dt, 0, ns,
                                ctab.data(), stab.data(), active.data()),
              pre);
    wf_ctx *ctx = nullptr;
    check(wf_create((uint32_t)c.nx, (uint32_t)c.ny, (uint32_t)c.nz, ids.data(), table.data(),
                    pml ? shell : nullptr, pml ? prof.data() : nullptr, 0, 0, &ctx),
          pre);
    std::vector<wf_source_site> ssites;
    for (const wf_source_desc &d : sources) ssites.push_back(d.site);
    double secs = 0.0;
    std::vector<double> sc(nf), ss(nf);
    std::vector<uint64_t> cnt(nprobe);
    try {
        check(wf_set_sources(ctx, (uint32_t)nsrc, ssites.data(), ns, values.data()), pre);
        check(wf_set_probes(ctx, (uint32_t)nprobe, psites.data(), ns, ctab.data(), stab.data(),
                            active.data(), 0),
              pre);
        check(wf_run(ctx, ns, every, c.output_path.c_str(), &secs), pre);
        check(wf_read_probe_sums(ctx, sc.data(), ss.data(), cnt.data()), pre);
    } catch (...) {
        wf_destroy(ctx);
        throw;
    }
    wf_destroy(ctx);
    std::fprintf(stderr, "wavefront: %llu steps over 1 slab(s) in %ss (%s steps/s)\n", (unsigned long long)ns,
                 rust_fixed(secs, 3).c_str(), rust_fixed((double)ns / std::max(secs, 1e-9), 1).c_str());

    uint64_t off = 0;
    for (uint64_t p = 0; p < nprobe; ++p) {
        const wf_probe_site &q = psites[p];
        std::fprintf(stderr, "wavefront: probe (%u, %u, %u) %s frequency response:\n", q.x, q.y, q.z,
                     component_name(q.component));
        std::vector<double> amp(q.nfreq), ph(q.nfreq);
        wf_probe_spectrum(q.nfreq, sc.data() + off, ss.data() + off, cnt[p], amp.data(), ph.data());
        for (uint32_t i = 0; i < q.nfreq; ++i)
            std::fprintf(stderr, "  %s Hz: amplitude %s, phase %s rad\n", rust_exp(freqs[off + i], 4).c_str(),
                         rust_exp(amp[i], 6).c_str(), rust_fixed(ph[i], 4).c_str());
        off += q.nfreq;
    }
}

}  // namespace

int main(int argc, char **argv)
{
    Config c;
    try {
        c = parse_args(argc, argv);
    } catch (const ParseError &e) {
        std::fprintf(stderr, "wavefront: %s\n\n", e.msg.c_str());
        print_usage();
        return 1;
    }
    try {
        run(c);
    } catch (const RunError &e) {
        std::fprintf(stderr, "wavefront: %s\n", e.msg.c_str());
        return e.exit_code;
    }
    return 0;
}

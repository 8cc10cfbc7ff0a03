// Timestep kernels for gfx950.
//
// Every lane owns four consecutive x-cells (one 16-byte vector per field
// component): 16-byte accesses sustain about 6.5 TB/s on MI355X where 4-byte
// accesses stall near 4 TB/s.  x-neighbours come from the adjacent lane
// (__shfl_down / __shfl_up of one component); the lane at a wave edge loads
// its one outside cell directly.
//
// Arithmetic: built with -ffp-contract=off; every f32 operation is a single
// rounded IEEE operation in the order of src/fdtd.rs, so the results are
// bit-identical to the CPU engine (vector operations act per element).
#include <hip/hip_runtime.h>

#include "wf_kernels.h"

namespace wf {
namespace {

__device__ __forceinline__ f4 ld4(const float *p) { return *reinterpret_cast<const f4 *>(p); }

__device__ __forceinline__ void st4(float *p, f4 v)
{
    __builtin_nontemporal_store(v, reinterpret_cast<f4 *>(p));
}

__device__ __forceinline__ f4 splat(float v) { return f4{v, v, v, v}; }

// Per-cell coefficient pair for the four material ids packed in `ids`.
__device__ __forceinline__ void coef4(const float2 *s, uint32_t ids, f4 &c0, f4 &c1)
{
    const float2 a = s[ids & 255u], b = s[(ids >> 8) & 255u], c = s[(ids >> 16) & 255u],
                 d = s[ids >> 24];
    c0 = f4{a.x, b.x, c.x, d.x};
    c1 = f4{a.y, b.y, c.y, d.y};
}

// Per-lane CPML description of a quad (the graded layers and the shell are
// whole 8-cell blocks, so every flag is uniform over a quad).
struct QuadPml {
    bool shell_xy, gx, gy;
    int gxi, gyi;
    f4 pxb, pxa, pxk;
    float pyb, pya, pyk;
};

__device__ __forceinline__ QuadPml quad_pml(const Geo &g, const Ptrs &p, int x, int y)
{
    QuadPml q;
    q.shell_xy = in_shell_axis(g, 0, x) || in_shell_axis(g, 1, y);
    q.gx = graded(g, 0, x, g.nx);
    q.gy = graded(g, 1, y, g.ny);
    q.gxi = slab_index(g, 0, x, g.nx);
    q.gyi = slab_index(g, 1, y, g.ny);
    q.pxb = ld4(p.prof[0][0] + x);
    q.pxa = ld4(p.prof[0][1] + x);
    q.pxk = ld4(p.prof[0][2] + x);
    q.pyb = p.prof[1][0][y];
    q.pya = p.prof[1][1][y];
    q.pyk = p.prof[1][2][y];
    return q;
}

// ===========================================================================
// Split variant: update_h_field[_pml] then update_e_field[_pml], in place.
// Workgroup = kSY waves; wave w owns row y, lane l the quad x = 4*(64*tx+l);
// each thread marches z over a chunk with E(z+1) / H(z-1) carried.
// ===========================================================================
constexpr int kSY = 4;

template <bool PML>
__global__ __launch_bounds__(64 * kSY) void k_h4(Geo g, Ptrs p, int zchunk, int gx, int gy, int gz)
{
    __shared__ float2 s_co[256];  // (da, db)
    const int lane = threadIdx.x, w = threadIdx.y;
    for (int i = w * 64 + lane; i < 256; i += 64 * kSY) {
        const float4 c = p.coef[i];
        s_co[i] = make_float2(c.z, c.w);
    }
    __syncthreads();

    const int t = xcd_tile(blockIdx.x, gx * gy * gz);
    const int tx = t % gx, ty = (t / gx) % gy, tz = t / (gx * gy);
    const int x = (tx * 64 + lane) * 4;
    const int y = ty * kSY + w;
    const bool live = x < g.nx && y < g.ny;
    const int z0 = tz * zchunk, z1 = min(z0 + zchunk, g.nz);
    const size_t nx = g.nx, plane = nx * g.ny;
    const int yc = live ? y : 0, xc = live ? x : 0;
    const size_t row = (size_t)yc * nx + xc;
    const int yp = plus1(yc, g.ny, g.periodic & 2u);
    const size_t rowp_y = (size_t)(yp >= 0 ? yp : 0) * nx + xc;
    const bool edge = lane == 63 || x + 4 >= g.nx;
    const int xq = plus1(xc + 3, g.nx, g.periodic & 1u);
    const size_t rowq = (size_t)yc * nx + (xq >= 0 ? xq : 0);

    float *__restrict__ ex = p.f[0];
    float *__restrict__ ey = p.f[1];
    float *__restrict__ ez = p.f[2];
    float *__restrict__ hx = p.f[3];
    float *__restrict__ hy = p.f[4];
    float *__restrict__ hz = p.f[5];

    QuadPml q{};
    if (PML && live) q = quad_pml(g, p, x, y);

    f4 ex_c = splat(0.f), ey_c = splat(0.f);
    if (live) {
        ex_c = ld4(ex + (size_t)z0 * plane + row);
        ey_c = ld4(ey + (size_t)z0 * plane + row);
    }
    for (int z = z0; z < z1; ++z) {
        const size_t zo = (size_t)z * plane;
        const int zp = plus1(z, g.nz, g.periodic & 4u);
        f4 ez_c = splat(0.f), ex_n = splat(0.f), ey_n = splat(0.f), ez_py = splat(0.f),
           ex_py = splat(0.f);
        if (live) {
            ez_c = ld4(ez + zo + row);
            if (zp >= 0) {
                ex_n = ld4(ex + (size_t)zp * plane + row);
                ey_n = ld4(ey + (size_t)zp * plane + row);
            }
            if (yp >= 0) {
                ez_py = ld4(ez + zo + rowp_y);
                ex_py = ld4(ex + zo + rowp_y);
            }
        }
        float sez = __shfl_down(ez_c.x, 1, 64), sey = __shfl_down(ey_c.x, 1, 64);
        if (edge) {
            sez = 0.f;
            sey = 0.f;
            if (live && xq >= 0) {
                sez = ez[zo + rowq];
                sey = ey[zo + rowq];
            }
        }
        const f4 ez_px = f4{ez_c.y, ez_c.z, ez_c.w, sez};
        const f4 ey_px = f4{ey_c.y, ey_c.z, ey_c.w, sey};

        f4 d_ez_dy = ez_py - ez_c, d_ex_dy = ex_py - ex_c;
        f4 d_ey_dz = ey_n - ey_c, d_ex_dz = ex_n - ex_c;
        f4 d_ez_dx = ez_px - ez_c, d_ey_dx = ey_px - ey_c;
        if (PML && live && (q.shell_xy || in_shell_axis(g, 2, z))) {
            if (q.gy) {
                const size_t i = ((size_t)z * (2 * g.T[1]) + q.gyi) * nx + x;
                f4 p0 = ld4(p.psi[1][0] + i), p1 = ld4(p.psi[1][1] + i);
                d_ez_dy = cpml4(d_ez_dy, splat(q.pyb), splat(q.pya), splat(q.pyk), p0);
                d_ex_dy = cpml4(d_ex_dy, splat(q.pyb), splat(q.pya), splat(q.pyk), p1);
                st4(p.psi[1][0] + i, p0);
                st4(p.psi[1][1] + i, p1);
            } else {
                d_ez_dy = d_ez_dy + 0.0f;
                d_ex_dy = d_ex_dy + 0.0f;
            }
            if (graded(g, 2, z, g.nz)) {
                const size_t i = ((size_t)slab_index(g, 2, z, g.nz) * g.ny + y) * nx + x;
                const f4 b = splat(p.prof[2][0][z]), a = splat(p.prof[2][1][z]),
                         ik = splat(p.prof[2][2][z]);
                f4 p0 = ld4(p.psi[2][0] + i), p1 = ld4(p.psi[2][1] + i);
                d_ey_dz = cpml4(d_ey_dz, b, a, ik, p0);
                d_ex_dz = cpml4(d_ex_dz, b, a, ik, p1);
                st4(p.psi[2][0] + i, p0);
                st4(p.psi[2][1] + i, p1);
            } else {
                d_ey_dz = d_ey_dz + 0.0f;
                d_ex_dz = d_ex_dz + 0.0f;
            }
            if (q.gx) {
                const size_t i = ((size_t)z * g.ny + y) * (2 * g.T[0]) + q.gxi;
                f4 p0 = ld4(p.psi[0][0] + i), p1 = ld4(p.psi[0][1] + i);
                d_ez_dx = cpml4(d_ez_dx, q.pxb, q.pxa, q.pxk, p0);
                d_ey_dx = cpml4(d_ey_dx, q.pxb, q.pxa, q.pxk, p1);
                st4(p.psi[0][0] + i, p0);
                st4(p.psi[0][1] + i, p1);
            } else {
                d_ez_dx = d_ez_dx + 0.0f;
                d_ey_dx = d_ey_dx + 0.0f;
            }
        }
        if (live) {
            f4 da, db;
            coef4(s_co, *reinterpret_cast<const uint32_t *>(p.id + zo + row), da, db);
            const f4 chx = d_ez_dy - d_ey_dz;
            const f4 chy = d_ex_dz - d_ez_dx;
            const f4 chz = d_ey_dx - d_ex_dy;
            st4(hx + zo + row, da * ld4(hx + zo + row) - db * chx);
            st4(hy + zo + row, da * ld4(hy + zo + row) - db * chy);
            st4(hz + zo + row, da * ld4(hz + zo + row) - db * chz);
        }
        ex_c = ex_n;
        ey_c = ey_n;
    }
}

template <bool PML>
__global__ __launch_bounds__(64 * kSY) void k_e4(Geo g, Ptrs p, int zchunk, int gx, int gy, int gz)
{
    __shared__ float2 s_co[256];  // (ca, cb)
    const int lane = threadIdx.x, w = threadIdx.y;
    for (int i = w * 64 + lane; i < 256; i += 64 * kSY) {
        const float4 c = p.coef[i];
        s_co[i] = make_float2(c.x, c.y);
    }
    __syncthreads();

    const int t = xcd_tile(blockIdx.x, gx * gy * gz);
    const int tx = t % gx, ty = (t / gx) % gy, tz = t / (gx * gy);
    const int x = (tx * 64 + lane) * 4;
    const int y = ty * kSY + w;
    const bool live = x < g.nx && y < g.ny;
    const int z0 = tz * zchunk, z1 = min(z0 + zchunk, g.nz);
    const size_t nx = g.nx, plane = nx * g.ny;
    const int yc = live ? y : 0, xc = live ? x : 0;
    const size_t row = (size_t)yc * nx + xc;
    const int ym = minus1(yc, g.ny, g.periodic & 2u);
    const size_t rowm_y = (size_t)(ym >= 0 ? ym : 0) * nx + xc;
    const bool edge = lane == 0 || x == 0;
    const int xq = minus1(xc, g.nx, g.periodic & 1u);
    const size_t rowq = (size_t)yc * nx + (xq >= 0 ? xq : 0);

    float *__restrict__ ex = p.f[0];
    float *__restrict__ ey = p.f[1];
    float *__restrict__ ez = p.f[2];
    float *__restrict__ hx = p.f[3];
    float *__restrict__ hy = p.f[4];
    float *__restrict__ hz = p.f[5];

    QuadPml q{};
    if (PML && live) q = quad_pml(g, p, x, y);

    f4 hx_mz = splat(0.f), hy_mz = splat(0.f);
    {
        const int zm = minus1(z0, g.nz, g.periodic & 4u);
        if (live && zm >= 0) {
            hx_mz = ld4(hx + (size_t)zm * plane + row);
            hy_mz = ld4(hy + (size_t)zm * plane + row);
        }
    }
    for (int z = z0; z < z1; ++z) {
        const size_t zo = (size_t)z * plane;
        f4 hx_c = splat(0.f), hy_c = splat(0.f), hz_c = splat(0.f), hz_my = splat(0.f),
           hx_my = splat(0.f);
        if (live) {
            hx_c = ld4(hx + zo + row);
            hy_c = ld4(hy + zo + row);
            hz_c = ld4(hz + zo + row);
            if (ym >= 0) {
                hz_my = ld4(hz + zo + rowm_y);
                hx_my = ld4(hx + zo + rowm_y);
            }
        }
        float shz = __shfl_up(hz_c.w, 1, 64), shy = __shfl_up(hy_c.w, 1, 64);
        if (edge) {
            shz = 0.f;
            shy = 0.f;
            if (live && xq >= 0) {
                shz = hz[zo + rowq];
                shy = hy[zo + rowq];
            }
        }
        const f4 hz_mx = f4{shz, hz_c.x, hz_c.y, hz_c.z};
        const f4 hy_mx = f4{shy, hy_c.x, hy_c.y, hy_c.z};

        f4 d_hz_dy = hz_c - hz_my, d_hx_dy = hx_c - hx_my;
        f4 d_hy_dz = hy_c - hy_mz, d_hx_dz = hx_c - hx_mz;
        f4 d_hz_dx = hz_c - hz_mx, d_hy_dx = hy_c - hy_mx;
        if (PML && live && (q.shell_xy || in_shell_axis(g, 2, z))) {
            if (q.gy) {
                const size_t i = ((size_t)z * (2 * g.T[1]) + q.gyi) * nx + x;
                f4 p0 = ld4(p.psi[1][2] + i), p1 = ld4(p.psi[1][3] + i);
                d_hz_dy = cpml4(d_hz_dy, splat(q.pyb), splat(q.pya), splat(q.pyk), p0);
                d_hx_dy = cpml4(d_hx_dy, splat(q.pyb), splat(q.pya), splat(q.pyk), p1);
                st4(p.psi[1][2] + i, p0);
                st4(p.psi[1][3] + i, p1);
            } else {
                d_hz_dy = d_hz_dy + 0.0f;
                d_hx_dy = d_hx_dy + 0.0f;
            }
            if (graded(g, 2, z, g.nz)) {
                const size_t i = ((size_t)slab_index(g, 2, z, g.nz) * g.ny + y) * nx + x;
                const f4 b = splat(p.prof[2][0][z]), a = splat(p.prof[2][1][z]),
                         ik = splat(p.prof[2][2][z]);
                f4 p0 = ld4(p.psi[2][2] + i), p1 = ld4(p.psi[2][3] + i);
                d_hy_dz = cpml4(d_hy_dz, b, a, ik, p0);
                d_hx_dz = cpml4(d_hx_dz, b, a, ik, p1);
                st4(p.psi[2][2] + i, p0);
                st4(p.psi[2][3] + i, p1);
            } else {
                d_hy_dz = d_hy_dz + 0.0f;
                d_hx_dz = d_hx_dz + 0.0f;
            }
            if (q.gx) {
                const size_t i = ((size_t)z * g.ny + y) * (2 * g.T[0]) + q.gxi;
                f4 p0 = ld4(p.psi[0][2] + i), p1 = ld4(p.psi[0][3] + i);
                d_hz_dx = cpml4(d_hz_dx, q.pxb, q.pxa, q.pxk, p0);
                d_hy_dx = cpml4(d_hy_dx, q.pxb, q.pxa, q.pxk, p1);
                st4(p.psi[0][2] + i, p0);
                st4(p.psi[0][3] + i, p1);
            } else {
                d_hz_dx = d_hz_dx + 0.0f;
                d_hy_dx = d_hy_dx + 0.0f;
            }
        }
        if (live) {
            f4 ca, cb;
            coef4(s_co, *reinterpret_cast<const uint32_t *>(p.id + zo + row), ca, cb);
            const f4 cex = d_hz_dy - d_hy_dz;
            const f4 cey = d_hx_dz - d_hz_dx;
            const f4 cez = d_hy_dx - d_hx_dy;
            st4(ex + zo + row, ca * ld4(ex + zo + row) + cb * cex);
            st4(ey + zo + row, ca * ld4(ey + zo + row) + cb * cey);
            st4(ez + zo + row, ca * ld4(ez + zo + row) + cb * cez);
        }
        hx_mz = hx_c;
        hy_mz = hy_c;
    }
}

// ===========================================================================
// Fused variant: H(n+1/2) and E(n+1) in one sweep, state A -> state B.
//
// Per cell 24 B read + 24 B written + 1 B id (the two-pass variant moves
// 74 B).  Workgroup = kFY waves x 64 lanes; wave w holds row
// y = ty*(kFY-1) - 1 + w, lane l the quad x = 256*tx + 4l.  Wave 0 is a
// one-row halo: it recomputes H(n+1/2) of row y0-1 from state A (the same
// arithmetic as that row's owner, so the same bits) for the E update of the
// owned rows, and never stores.  H at x-1 comes from the previous lane; lane
// 0 of every owned row also recomputes the one halo cell x0-1.  The sweep
// marches z over a chunk, starting one plane early for H(z0-1); E(z+1) and
// H(z-1) ride in registers; H at y-1 crosses waves through a double-buffered
// LDS plane behind an LDS-only barrier, so stores are not waited for.
// ===========================================================================
constexpr int kFY = 8;

// H(n+1/2) of the single cell (x, y, z) on the x-halo, hy and hz only (the E
// update at x+1 reads no other component).  e_* are E^n at (x, y, z); exn,
// eyn at (x, y, z+1); ex_py at (x, y+1, z); ez_px, ey_px at (x+1, y, z).
template <bool PML>
__device__ __forceinline__ void halo_h(const Geo &g, const Ptrs &a, const float2 *s_dh, int x,
                                       int y, int z, float ex_c, float ey_c, float ez_c,
                                       float exn, float eyn, float ex_py, float ez_px,
                                       float ey_px, int mid, float hy_o, float hz_o, float &hyn,
                                       float &hzn)
{
    float d_ex_dy = ex_py - ex_c;
    float d_ex_dz = exn - ex_c;
    float d_ez_dx = ez_px - ez_c, d_ey_dx = ey_px - ey_c;
    (void)eyn;
    if (PML && (in_shell_axis(g, 0, x) || in_shell_axis(g, 1, y) || in_shell_axis(g, 2, z))) {
        const size_t nx = g.nx;
        if (graded(g, 1, y, g.ny)) {
            const size_t i = ((size_t)z * (2 * g.T[1]) + slab_index(g, 1, y, g.ny)) * nx + x;
            float p1 = a.psi[1][1][i];
            d_ex_dy = cpml(d_ex_dy, a.prof[1][0][y], a.prof[1][1][y], a.prof[1][2][y], p1);
        } else {
            d_ex_dy = d_ex_dy + 0.0f;
        }
        if (graded(g, 2, z, g.nz)) {
            const size_t i = ((size_t)slab_index(g, 2, z, g.nz) * g.ny + y) * nx + x;
            float p1 = a.psi[2][1][i];
            d_ex_dz = cpml(d_ex_dz, a.prof[2][0][z], a.prof[2][1][z], a.prof[2][2][z], p1);
        } else {
            d_ex_dz = d_ex_dz + 0.0f;
        }
        if (graded(g, 0, x, g.nx)) {
            const size_t i = ((size_t)z * g.ny + y) * (2 * g.T[0]) + slab_index(g, 0, x, g.nx);
            float p0 = a.psi[0][0][i], p1 = a.psi[0][1][i];
            const float b = a.prof[0][0][x], aa = a.prof[0][1][x], ik = a.prof[0][2][x];
            d_ez_dx = cpml(d_ez_dx, b, aa, ik, p0);
            d_ey_dx = cpml(d_ey_dx, b, aa, ik, p1);
        } else {
            d_ez_dx = d_ez_dx + 0.0f;
            d_ey_dx = d_ey_dx + 0.0f;
        }
    }
    const float2 c = s_dh[mid];
    hyn = c.x * hy_o - c.y * (d_ex_dz - d_ez_dx);
    hzn = c.x * hz_o - c.y * (d_ey_dx - d_ex_dy);
}

template <bool PML>
__global__ __launch_bounds__(64 * kFY) void k_fused4(Geo g, Ptrs a, Ptrs b, int zchunk, int gx,
                                                     int gy, int gz)
{
    __shared__ float2 s_ce[256];          // (ca, cb)
    __shared__ float2 s_ch[256];          // (da, db)
    __shared__ f4 s_h[2][2][kFY][64];     // [buf][hz, hx][wave][lane]
    const int lane = threadIdx.x, w = threadIdx.y;
    for (int i = w * 64 + lane; i < 256; i += 64 * kFY) {
        const float4 c = a.coef[i];
        s_ce[i] = make_float2(c.x, c.y);
        s_ch[i] = make_float2(c.z, c.w);
    }
    __syncthreads();

    const int t = xcd_tile(blockIdx.x, gx * gy * gz);
    const int tx = t % gx, ty = (t / gx) % gy, tz = t / (gx * gy);
    const int x = tx * 256 + lane * 4;
    const int yr = ty * (kFY - 1) - 1 + w;
    int y = yr;
    bool live = x < g.nx && yr < g.ny;
    if (yr < 0) {
        if (g.periodic & 2u) y = g.ny - 1; else live = false;
    }
    const bool own = live && w >= 1;
    const int xc = live ? x : 0, yc = live ? y : 0;
    const size_t nx = g.nx, plane = nx * g.ny;
    const size_t row = (size_t)yc * nx + xc;
    const int yp = plus1(yc, g.ny, g.periodic & 2u);
    const size_t rowp_y = (size_t)(yp >= 0 ? yp : 0) * nx + xc;
    const bool edge = lane == 63 || x + 4 >= g.nx;
    const int xq = plus1(xc + 3, g.nx, g.periodic & 1u);
    const size_t rowq = (size_t)yc * nx + (xq >= 0 ? xq : 0);
    // x-halo cell for lane 0 of owned rows
    const bool hcell = own && lane == 0;
    const int xh = minus1(xc, g.nx, g.periodic & 1u);  // -1: zero boundary
    const bool hlive = hcell && xh >= 0;
    const size_t rowh = (size_t)yc * nx + (xh >= 0 ? xh : 0);
    const size_t rowh_py = (size_t)(yp >= 0 ? yp : 0) * nx + (xh >= 0 ? xh : 0);

    const float *__restrict__ aex = a.f[0];
    const float *__restrict__ aey = a.f[1];
    const float *__restrict__ aez = a.f[2];
    const float *__restrict__ ahx = a.f[3];
    const float *__restrict__ ahy = a.f[4];
    const float *__restrict__ ahz = a.f[5];

    QuadPml q{};
    if (PML && live) q = quad_pml(g, a, x, y);

    const int z0 = tz * zchunk, z1 = min(z0 + zchunk, g.nz);
    const int zpre = minus1(z0, g.nz, g.periodic & 4u);
    const int it0 = zpre >= 0 ? -1 : 0;
    const int nit = z1 - z0;

    // E^n of the own quad (and of the halo cell) at the current plane;
    // H(n+1/2) of the own quad one plane below.
    f4 ex_c = splat(0.f), ey_c = splat(0.f), ez_c = splat(0.f);
    f4 hx_mz = splat(0.f), hy_mz = splat(0.f);
    float hex = 0.f, hey = 0.f, hez = 0.f;
    {
        const int zc = it0 < 0 ? zpre : z0;
        const size_t o = (size_t)zc * plane;
        if (live) {
            ex_c = ld4(aex + o + row);
            ey_c = ld4(aey + o + row);
            ez_c = ld4(aez + o + row);
        }
        if (hlive) {
            hex = aex[o + rowh];
            hey = aey[o + rowh];
            hez = aez[o + rowh];
        }
    }

    for (int it = it0; it < nit; ++it) {
        const int z = it < 0 ? zpre : z0 + it;
        const size_t zo = (size_t)z * plane;
        const int zp = plus1(z, g.nz, g.periodic & 4u);
        const bool store = own && it >= 0;

        // ---- loads (state A) ------------------------------------------------
        f4 ex_n = splat(0.f), ey_n = splat(0.f), ez_n = splat(0.f);
        f4 ez_py = splat(0.f), ex_py = splat(0.f);
        f4 hx_o = splat(0.f), hy_o = splat(0.f), hz_o = splat(0.f);
        uint32_t ids = 0;
        if (live) {
            if (zp >= 0) {
                const size_t on = (size_t)zp * plane + row;
                ex_n = ld4(aex + on);
                ey_n = ld4(aey + on);
                ez_n = ld4(aez + on);
            }
            if (yp >= 0) {
                ez_py = ld4(aez + zo + rowp_y);
                ex_py = ld4(aex + zo + rowp_y);
            }
            hx_o = ld4(ahx + zo + row);
            hy_o = ld4(ahy + zo + row);
            hz_o = ld4(ahz + zo + row);
            ids = *reinterpret_cast<const uint32_t *>(a.id + zo + row);
        }
        float sez = __shfl_down(ez_c.x, 1, 64), sey = __shfl_down(ey_c.x, 1, 64);
        if (edge) {
            sez = 0.f;
            sey = 0.f;
            if (live && xq >= 0) {
                sez = aez[zo + rowq];
                sey = aey[zo + rowq];
            }
        }

        // ---- H(n+1/2) of the own quad: update_h_field[_pml] -----------------
        const f4 ez_px = f4{ez_c.y, ez_c.z, ez_c.w, sez};
        const f4 ey_px = f4{ey_c.y, ey_c.z, ey_c.w, sey};
        f4 d_ez_dy = ez_py - ez_c, d_ex_dy = ex_py - ex_c;
        f4 d_ey_dz = ey_n - ey_c, d_ex_dz = ex_n - ex_c;
        f4 d_ez_dx = ez_px - ez_c, d_ey_dx = ey_px - ey_c;
        const bool shell = PML && live && (q.shell_xy || in_shell_axis(g, 2, z));
        const bool gzz = PML && graded(g, 2, z, g.nz);
        size_t ipy = 0, ipz = 0, ipx = 0;
        f4 pzb = splat(1.f), pza = splat(0.f), pzk = splat(1.f);
        if (PML && shell) {
            ipy = ((size_t)z * (2 * g.T[1]) + q.gyi) * nx + x;
            ipx = ((size_t)z * g.ny + y) * (2 * g.T[0]) + q.gxi;
            if (gzz) {
                ipz = ((size_t)slab_index(g, 2, z, g.nz) * g.ny + y) * nx + x;
                pzb = splat(a.prof[2][0][z]);
                pza = splat(a.prof[2][1][z]);
                pzk = splat(a.prof[2][2][z]);
            }
            if (q.gy) {
                f4 p0 = ld4(a.psi[1][0] + ipy), p1 = ld4(a.psi[1][1] + ipy);
                d_ez_dy = cpml4(d_ez_dy, splat(q.pyb), splat(q.pya), splat(q.pyk), p0);
                d_ex_dy = cpml4(d_ex_dy, splat(q.pyb), splat(q.pya), splat(q.pyk), p1);
                if (store) {
                    st4(b.psi[1][0] + ipy, p0);
                    st4(b.psi[1][1] + ipy, p1);
                }
            } else {
                d_ez_dy = d_ez_dy + 0.0f;
                d_ex_dy = d_ex_dy + 0.0f;
            }
            if (gzz) {
                f4 p0 = ld4(a.psi[2][0] + ipz), p1 = ld4(a.psi[2][1] + ipz);
                d_ey_dz = cpml4(d_ey_dz, pzb, pza, pzk, p0);
                d_ex_dz = cpml4(d_ex_dz, pzb, pza, pzk, p1);
                if (store) {
                    st4(b.psi[2][0] + ipz, p0);
                    st4(b.psi[2][1] + ipz, p1);
                }
            } else {
                d_ey_dz = d_ey_dz + 0.0f;
                d_ex_dz = d_ex_dz + 0.0f;
            }
            if (q.gx) {
                f4 p0 = ld4(a.psi[0][0] + ipx), p1 = ld4(a.psi[0][1] + ipx);
                d_ez_dx = cpml4(d_ez_dx, q.pxb, q.pxa, q.pxk, p0);
                d_ey_dx = cpml4(d_ey_dx, q.pxb, q.pxa, q.pxk, p1);
                if (store) {
                    st4(b.psi[0][0] + ipx, p0);
                    st4(b.psi[0][1] + ipx, p1);
                }
            } else {
                d_ez_dx = d_ez_dx + 0.0f;
                d_ey_dx = d_ey_dx + 0.0f;
            }
        }
        f4 da, db;
        coef4(s_ch, ids, da, db);
        f4 hxn = da * hx_o - db * (d_ez_dy - d_ey_dz);
        f4 hyn = da * hy_o - db * (d_ex_dz - d_ez_dx);
        f4 hzn = da * hz_o - db * (d_ey_dx - d_ex_dy);
        if (!live) {
            hxn = splat(0.f);
            hyn = splat(0.f);
            hzn = splat(0.f);
        }

        // ---- H(n+1/2) of the x-halo cell (lane 0 of owned rows) -------------
        float hy_h = 0.f, hz_h = 0.f;
        float hexn = 0.f, heyn = 0.f, hezn = 0.f;
        if (hlive) {
            if (zp >= 0) {
                const size_t on = (size_t)zp * plane + rowh;
                hexn = aex[on];
                heyn = aey[on];
                hezn = aez[on];
            }
            const float hex_py = yp >= 0 ? aex[zo + rowh_py] : 0.f;
            if (it >= 0)
                halo_h<PML>(g, a, s_ch, xh, y, z, hex, hey, hez, hexn, heyn, hex_py, ez_c.x,
                            ey_c.x, a.id[zo + rowh], ahy[zo + rowh], ahz[zo + rowh], hy_h, hz_h);
        }

        if (it >= 0) {
            // ---- neighbours of H(n+1/2) at x-1 (lane-1) and y-1 (LDS) -------
            float hz_l = __shfl_up(hzn.w, 1, 64), hy_l = __shfl_up(hyn.w, 1, 64);
            if (lane == 0) {
                hz_l = hz_h;
                hy_l = hy_h;
            }
            const int buf = it & 1;
            s_h[buf][0][w][lane] = hzn;
            s_h[buf][1][w][lane] = hxn;
            lds_barrier();
            if (store) {
                const f4 hz_my = s_h[buf][0][w - 1][lane];
                const f4 hx_my = s_h[buf][1][w - 1][lane];
                const f4 hz_mx = f4{hz_l, hzn.x, hzn.y, hzn.z};
                const f4 hy_mx = f4{hy_l, hyn.x, hyn.y, hyn.z};

                // ---- E(n+1): update_e_field[_pml] ---------------------------
                f4 d_hz_dy = hzn - hz_my, d_hx_dy = hxn - hx_my;
                f4 d_hy_dz = hyn - hy_mz, d_hx_dz = hxn - hx_mz;
                f4 d_hz_dx = hzn - hz_mx, d_hy_dx = hyn - hy_mx;
                if (PML && shell) {
                    if (q.gy) {
                        f4 p0 = ld4(a.psi[1][2] + ipy), p1 = ld4(a.psi[1][3] + ipy);
                        d_hz_dy = cpml4(d_hz_dy, splat(q.pyb), splat(q.pya), splat(q.pyk), p0);
                        d_hx_dy = cpml4(d_hx_dy, splat(q.pyb), splat(q.pya), splat(q.pyk), p1);
                        st4(b.psi[1][2] + ipy, p0);
                        st4(b.psi[1][3] + ipy, p1);
                    } else {
                        d_hz_dy = d_hz_dy + 0.0f;
                        d_hx_dy = d_hx_dy + 0.0f;
                    }
                    if (gzz) {
                        f4 p0 = ld4(a.psi[2][2] + ipz), p1 = ld4(a.psi[2][3] + ipz);
                        d_hy_dz = cpml4(d_hy_dz, pzb, pza, pzk, p0);
                        d_hx_dz = cpml4(d_hx_dz, pzb, pza, pzk, p1);
                        st4(b.psi[2][2] + ipz, p0);
                        st4(b.psi[2][3] + ipz, p1);
                    } else {
                        d_hy_dz = d_hy_dz + 0.0f;
                        d_hx_dz = d_hx_dz + 0.0f;
                    }
                    if (q.gx) {
                        f4 p0 = ld4(a.psi[0][2] + ipx), p1 = ld4(a.psi[0][3] + ipx);
                        d_hz_dx = cpml4(d_hz_dx, q.pxb, q.pxa, q.pxk, p0);
                        d_hy_dx = cpml4(d_hy_dx, q.pxb, q.pxa, q.pxk, p1);
                        st4(b.psi[0][2] + ipx, p0);
                        st4(b.psi[0][3] + ipx, p1);
                    } else {
                        d_hz_dx = d_hz_dx + 0.0f;
                        d_hy_dx = d_hy_dx + 0.0f;
                    }
                }
                f4 ca, cb;
                coef4(s_ce, ids, ca, cb);
                const size_t o = zo + row;
                st4(b.f[0] + o, ca * ex_c + cb * (d_hz_dy - d_hy_dz));
                st4(b.f[1] + o, ca * ey_c + cb * (d_hx_dz - d_hz_dx));
                st4(b.f[2] + o, ca * ez_c + cb * (d_hy_dx - d_hx_dy));
                st4(b.f[3] + o, hxn);
                st4(b.f[4] + o, hyn);
                st4(b.f[5] + o, hzn);
            }
        }
        hx_mz = hxn;
        hy_mz = hyn;
        ex_c = ex_n;
        ey_c = ey_n;
        ez_c = ez_n;
        hex = hexn;
        hey = heyn;
        hez = hezn;
    }
}

}  // namespace

hipError_t launch_split(const StepLaunch &s, hipEvent_t *ev)
{
    const int gx = (s.g.nx + 255) / 256, gy = (s.g.ny + kSY - 1) / kSY,
              gz = (s.g.nz + s.zchunk - 1) / s.zchunk;
    const dim3 grid((unsigned)(gx * gy * gz)), block(64, kSY);
    if (ev) (void)hipEventRecord(ev[0], s.stream);
    if (s.pml)
        hipLaunchKernelGGL(k_h4<true>, grid, block, 0, s.stream, s.g, s.cur, s.zchunk, gx, gy, gz);
    else
        hipLaunchKernelGGL(k_h4<false>, grid, block, 0, s.stream, s.g, s.cur, s.zchunk, gx, gy, gz);
    if (ev) (void)hipEventRecord(ev[1], s.stream);
    if (ev) (void)hipEventRecord(ev[2], s.stream);
    if (s.pml)
        hipLaunchKernelGGL(k_e4<true>, grid, block, 0, s.stream, s.g, s.cur, s.zchunk, gx, gy, gz);
    else
        hipLaunchKernelGGL(k_e4<false>, grid, block, 0, s.stream, s.g, s.cur, s.zchunk, gx, gy, gz);
    if (ev) (void)hipEventRecord(ev[3], s.stream);
    return hipGetLastError();
}

hipError_t launch_fused(const StepLaunch &s, hipEvent_t *ev)
{
    const int gx = (s.g.nx + 255) / 256, gy = (s.g.ny + kFY - 2) / (kFY - 1),
              gz = (s.g.nz + s.zchunk - 1) / s.zchunk;
    const dim3 grid((unsigned)(gx * gy * gz)), block(64, kFY);
    if (ev) (void)hipEventRecord(ev[0], s.stream);
    if (s.pml)
        hipLaunchKernelGGL(k_fused4<true>, grid, block, 0, s.stream, s.g, s.cur, s.alt, s.zchunk,
                           gx, gy, gz);
    else
        hipLaunchKernelGGL(k_fused4<false>, grid, block, 0, s.stream, s.g, s.cur, s.alt,
                           s.zchunk, gx, gy, gz);
    if (ev) (void)hipEventRecord(ev[1], s.stream);
    if (ev) (void)hipEventRecord(ev[2], s.stream);
    if (ev) (void)hipEventRecord(ev[3], s.stream);
    return hipGetLastError();
}

}  // namespace wf

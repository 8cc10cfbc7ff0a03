// Timestep kernels for gfx950.
//
// Every lane owns four consecutive x-cells (one 16-byte vector per field
// component): 16-byte accesses sustain about 6.5 TB/s on MI355X where 4-byte
// accesses stall near 4 TB/s.  x-neighbours come from the adjacent lane
// (__shfl_down / __shfl_up of one component); a lane at a segment edge loads
// its one outside cell directly.
//
// Work is described by boxes of cells.  The fused single-pass kernel covers
// the interior box, whose cells and one-cell halos are all outside the CPML
// shell; the split H and E kernels cover everything else (the "frame" that
// holds the whole CPML shell), or the whole domain in the split variant.
//
// Arithmetic: built with -ffp-contract=off; every f32 operation is a single
// rounded IEEE operation in the order of src/fdtd.rs, so the results are
// bit-identical to the CPU engine (vector operations act per element).
// The same file is built a second time with WF_STEP_FMA and
// -ffp-contract=fast: the contracted kernel set (namespace wf::fma), whose
// a*b+c become fused multiply-adds (results within rounding of the exact
// ones, no longer bit-identical).
#include <hip/hip_runtime.h>

#include "wf_kernels.h"

namespace wf {
#ifdef WF_STEP_FMA
namespace fma {
#endif
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

// Locates workgroup `i` (already XCD-renumbered) in the box table.
__device__ __forceinline__ const Box &find_box(const BoxList &bl, int i, int &local)
{
    int k = 0;
    while (k + 1 < bl.n && i >= bl.b[k + 1].first) ++k;
    local = i - bl.b[k].first;
    return bl.b[k];
}

// ===========================================================================
// Split kernels: update_h_field[_pml] over the boxes, then
// update_e_field[_pml] over the same boxes.
//
// A wave is rpw = 64/lpr (rounded down) rows of lpr lanes, lpr any width up
// to 64 (lanes past rpw*lpr idle); lane (r, l) owns the quad
// x = x0 + 4*(lpr*tx + l) of row y = y0 + (kSY*rpw)*ty + w*rpw + r.
// Each thread marches z over a chunk with E(z+1) (H kernel) or H(z-1)
// (E kernel) carried in registers.  Sources and destinations are separate so
// the same kernels serve the in-place split variant (src == dst) and the
// frame of the fused variant (state A -> state B).
// ===========================================================================
constexpr int kSY = 4;

struct SplitGeo {
    int x, y, z0, z1;
    bool live, edge;
};

__device__ __forceinline__ SplitGeo split_geo(const Box &b, int local, int zchunk, bool hphase)
{
    const int lane = threadIdx.x, w = threadIdx.y;
    const int rpw = 64 / b.lpr;
    const int r = lane / b.lpr, l = lane % b.lpr;
    const int tx = local % b.gx, ty = (local / b.gx) % b.gy, tz = local / (b.gx * b.gy);
    SplitGeo s;
    s.x = b.x0 + 4 * (b.lpr * tx + l);
    s.y = b.y0 + (kSY * rpw) * ty + w * rpw + r;
    s.z0 = b.z0 + tz * zchunk;
    s.z1 = min(s.z0 + zchunk, b.z1);
    s.live = s.x < b.x1 && s.y < b.y1 && r < rpw;
    s.edge = hphase ? (l == b.lpr - 1 || s.x + 4 >= b.x1) : (l == 0 || s.x == b.x0);
    return s;
}

// One workgroup of the H kernel: tile `local` of box b.  s_co: LDS for the
// (da, db) table.
template <bool PML, bool GHOST = false>
__device__ __forceinline__ void h4_tile(const Geo &g, const Ptrs &src, const Ptrs &dst,
                                        const Box &b, int local, int zchunk, float2 *s_co)
{
    const int lane = threadIdx.x, w = threadIdx.y;
    for (int i = w * 64 + lane; i < 256; i += 64 * kSY) {
        const float4 c = src.coef[i];
        s_co[i] = make_float2(c.z, c.w);
    }
    __syncthreads();

    const SplitGeo s = split_geo(b, local, zchunk, true);
    const bool live = s.live;
    const int x = live ? s.x : 0, y = live ? s.y : 0;
    const size_t nx = g.nx, plane = nx * g.ny;
    const size_t row = (size_t)y * nx + x;
    const int yp = plus1(y, g.ny, g.periodic & 2u);
    const size_t rowp_y = (size_t)(yp >= 0 ? yp : 0) * nx + x;
    const int xq = plus1(x + 3, g.nx, g.periodic & 1u);
    const size_t rowq = (size_t)y * nx + (xq >= 0 ? xq : 0);

    const float *__restrict__ ex = src.f[0];
    const float *__restrict__ ey = src.f[1];
    const float *__restrict__ ez = src.f[2];

    bool shell_xy = false, gx = false, gy = false;
    int gxi = 0, gyi = 0;
    if (PML && live) {
        shell_xy = in_shell_axis(g, 0, x) || in_shell_axis(g, 1, y);
        gx = graded(g, 0, x, g.nx);
        gy = graded(g, 1, y, g.ny);
        gxi = slab_index(g, 0, x, g.nx);
        gyi = slab_index(g, 1, y, g.ny);
    }

    f4 ex_c = splat(0.f), ey_c = splat(0.f);
    if (live) {
        ex_c = ld4(ex + (size_t)s.z0 * plane + row);
        ey_c = ld4(ey + (size_t)s.z0 * plane + row);
    }
    for (int z = s.z0; z < s.z1; ++z) {
        const size_t zo = (size_t)z * plane;
        const int zp = plus1(z, g.nz, g.periodic & 4u);
        f4 ez_c = splat(0.f), ex_n = splat(0.f), ey_n = splat(0.f), ez_py = splat(0.f),
           ex_py = splat(0.f);
        if (live) {
            ez_c = ld4(ez + zo + row);
            if (zp >= 0) {
                ex_n = ld4(ex + (size_t)zp * plane + row);
                ey_n = ld4(ey + (size_t)zp * plane + row);
            } else if (GHOST && src.ghost[5][0]) {
                ex_n = ld4(src.ghost[5][0] + (size_t)y * nx + x);
                ey_n = ld4(src.ghost[5][1] + (size_t)y * nx + x);
            }
            if (yp >= 0) {
                ez_py = ld4(ez + zo + rowp_y);
                ex_py = ld4(ex + zo + rowp_y);
            } else if (GHOST && src.ghost[3][0]) {
                ez_py = ld4(src.ghost[3][0] + (size_t)z * nx + x);
                ex_py = ld4(src.ghost[3][1] + (size_t)z * nx + x);
            }
        }
        // the next lane's cell; a row's last lane loads its own (edge)
        float sez = __shfl_down(ez_c.x, 1, 64), sey = __shfl_down(ey_c.x, 1, 64);
        if (s.edge) {
            sez = 0.f;
            sey = 0.f;
            if (live && xq >= 0) {
                sez = ez[zo + rowq];
                sey = ey[zo + rowq];
            } else if (GHOST && live && src.ghost[1][0]) {
                sey = src.ghost[1][0][(size_t)z * g.ny + y];
                sez = src.ghost[1][1][(size_t)z * g.ny + y];
            }
        }
        const f4 ez_px = f4{ez_c.y, ez_c.z, ez_c.w, sez};
        const f4 ey_px = f4{ey_c.y, ey_c.z, ey_c.w, sey};

        f4 d_ez_dy = ez_py - ez_c, d_ex_dy = ex_py - ex_c;
        f4 d_ey_dz = ey_n - ey_c, d_ex_dz = ex_n - ex_c;
        f4 d_ez_dx = ez_px - ez_c, d_ey_dx = ey_px - ey_c;
        if (PML && live && (shell_xy || in_shell_axis(g, 2, z))) {
            if (gy) {
                const size_t i = ((size_t)z * (2 * g.T[1]) + gyi) * nx + x;
                const f4 pb = splat(src.prof[1][0][y]), pa = splat(src.prof[1][1][y]),
                         pk = splat(src.prof[1][2][y]);
                f4 p0 = ld4(src.psi[1][0] + i), p1 = ld4(src.psi[1][1] + i);
                d_ez_dy = cpml4(d_ez_dy, pb, pa, pk, p0);
                d_ex_dy = cpml4(d_ex_dy, pb, pa, pk, p1);
                st4(dst.psi[1][0] + i, p0);
                st4(dst.psi[1][1] + i, p1);
            } else {
                d_ez_dy = d_ez_dy + 0.0f;
                d_ex_dy = d_ex_dy + 0.0f;
            }
            if (graded(g, 2, z, g.nz)) {
                const size_t i = ((size_t)slab_index(g, 2, z, g.nz) * g.ny + y) * nx + x;
                const f4 pb = splat(src.prof[2][0][z]), pa = splat(src.prof[2][1][z]),
                         pk = splat(src.prof[2][2][z]);
                f4 p0 = ld4(src.psi[2][0] + i), p1 = ld4(src.psi[2][1] + i);
                d_ey_dz = cpml4(d_ey_dz, pb, pa, pk, p0);
                d_ex_dz = cpml4(d_ex_dz, pb, pa, pk, p1);
                st4(dst.psi[2][0] + i, p0);
                st4(dst.psi[2][1] + i, p1);
            } else {
                d_ey_dz = d_ey_dz + 0.0f;
                d_ex_dz = d_ex_dz + 0.0f;
            }
            if (gx) {
                const size_t i = ((size_t)z * g.ny + y) * (2 * g.T[0]) + gxi;
                const f4 pb = ld4(src.prof[0][0] + x), pa = ld4(src.prof[0][1] + x),
                         pk = ld4(src.prof[0][2] + x);
                f4 p0 = ld4(src.psi[0][0] + i), p1 = ld4(src.psi[0][1] + i);
                d_ez_dx = cpml4(d_ez_dx, pb, pa, pk, p0);
                d_ey_dx = cpml4(d_ey_dx, pb, pa, pk, p1);
                st4(dst.psi[0][0] + i, p0);
                st4(dst.psi[0][1] + i, p1);
            } else {
                d_ez_dx = d_ez_dx + 0.0f;
                d_ey_dx = d_ey_dx + 0.0f;
            }
        }
        if (live) {
            f4 da, db;
            coef4(s_co, *reinterpret_cast<const uint32_t *>(src.id + zo + row), da, db);
            const f4 chx = d_ez_dy - d_ey_dz;
            const f4 chy = d_ex_dz - d_ez_dx;
            const f4 chz = d_ey_dx - d_ex_dy;
            st4(dst.f[3] + zo + row, da * ld4(src.f[3] + zo + row) - db * chx);
            st4(dst.f[4] + zo + row, da * ld4(src.f[4] + zo + row) - db * chy);
            st4(dst.f[5] + zo + row, da * ld4(src.f[5] + zo + row) - db * chz);
        }
        ex_c = ex_n;
        ey_c = ey_n;
    }
}

template <bool PML, bool GHOST>
__global__ __launch_bounds__(64 * kSY) void k_h4(Geo g, Ptrs src, Ptrs dst, BoxList bl,
                                                 int zchunk)
{
    __shared__ float2 s_co[256];  // (da, db)
    int local;
    const Box &b = find_box(bl, xcd_tile(blockIdx.x, bl.total), local);
    h4_tile<PML, GHOST>(g, src, dst, b, local, zchunk, s_co);
}

// Batched H kernel: every scene in place, one launch.  A grid smaller than
// the tile count makes it persistent: workgroup i takes tiles i, i + grid, ...
template <bool PML>
__global__ __launch_bounds__(64 * kSY) void k_h4_batch(const BatchScene *__restrict__ sc,
                                                       const uint32_t *__restrict__ wg_scene,
                                                       int total, int zchunk)
{
    __shared__ float2 s_co[256];  // (da, db)
    for (int i = blockIdx.x; i < total; i += gridDim.x) {
        const int t = xcd_tile(i, total);
        const BatchScene *d = sc + wg_scene[t];
        const Geo g = d->g;
        const Ptrs p = d->p;
        const Box b = d->box;
        if (t - b.first >= b.gx * b.gy * b.gz) return;  // tables and grid always agree
        h4_tile<PML>(g, p, p, b, t - b.first, zchunk, s_co);
        __syncthreads();  // the next tile reloads s_co
    }
}

// One workgroup of the E kernel.  hsrc: where H(n+1/2) lives; esrc: E^n and
// the psi memory of step n; dst: E(n+1) and psi of step n+1.
template <bool PML, bool GHOST = false>
__device__ __forceinline__ void e4_tile(const Geo &g, const Ptrs &hsrc, const Ptrs &esrc,
                                        const Ptrs &dst, const Box &b, int local, int zchunk,
                                        float2 *s_co)
{
    const int lane = threadIdx.x, w = threadIdx.y;
    for (int i = w * 64 + lane; i < 256; i += 64 * kSY) {
        const float4 c = esrc.coef[i];
        s_co[i] = make_float2(c.x, c.y);
    }
    __syncthreads();

    const SplitGeo s = split_geo(b, local, zchunk, false);
    const bool live = s.live;
    const int x = live ? s.x : 0, y = live ? s.y : 0;
    const size_t nx = g.nx, plane = nx * g.ny;
    const size_t row = (size_t)y * nx + x;
    const int ym = minus1(y, g.ny, g.periodic & 2u);
    const size_t rowm_y = (size_t)(ym >= 0 ? ym : 0) * nx + x;
    const int xq = minus1(x, g.nx, g.periodic & 1u);
    const size_t rowq = (size_t)y * nx + (xq >= 0 ? xq : 0);

    const float *__restrict__ hx = hsrc.f[3];
    const float *__restrict__ hy = hsrc.f[4];
    const float *__restrict__ hz = hsrc.f[5];

    bool shell_xy = false, gx = false, gy = false;
    int gxi = 0, gyi = 0;
    if (PML && live) {
        shell_xy = in_shell_axis(g, 0, x) || in_shell_axis(g, 1, y);
        gx = graded(g, 0, x, g.nx);
        gy = graded(g, 1, y, g.ny);
        gxi = slab_index(g, 0, x, g.nx);
        gyi = slab_index(g, 1, y, g.ny);
    }

    f4 hx_mz = splat(0.f), hy_mz = splat(0.f);
    {
        const int zm = minus1(s.z0, g.nz, g.periodic & 4u);
        if (live && zm >= 0) {
            hx_mz = ld4(hx + (size_t)zm * plane + row);
            hy_mz = ld4(hy + (size_t)zm * plane + row);
        } else if (GHOST && live && esrc.ghost[4][0]) {
            hx_mz = ld4(esrc.ghost[4][0] + (size_t)y * nx + x);
            hy_mz = ld4(esrc.ghost[4][1] + (size_t)y * nx + x);
        }
    }
    for (int z = s.z0; z < s.z1; ++z) {
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
            } else if (GHOST && esrc.ghost[2][0]) {
                hz_my = ld4(esrc.ghost[2][0] + (size_t)z * nx + x);
                hx_my = ld4(esrc.ghost[2][1] + (size_t)z * nx + x);
            }
        }
        float shz = __shfl_up(hz_c.w, 1, 64), shy = __shfl_up(hy_c.w, 1, 64);
        if (s.edge) {
            shz = 0.f;
            shy = 0.f;
            if (live && xq >= 0) {
                shz = hz[zo + rowq];
                shy = hy[zo + rowq];
            } else if (GHOST && live && esrc.ghost[0][0]) {
                shy = esrc.ghost[0][0][(size_t)z * g.ny + y];
                shz = esrc.ghost[0][1][(size_t)z * g.ny + y];
            }
        }
        const f4 hz_mx = f4{shz, hz_c.x, hz_c.y, hz_c.z};
        const f4 hy_mx = f4{shy, hy_c.x, hy_c.y, hy_c.z};

        f4 d_hz_dy = hz_c - hz_my, d_hx_dy = hx_c - hx_my;
        f4 d_hy_dz = hy_c - hy_mz, d_hx_dz = hx_c - hx_mz;
        f4 d_hz_dx = hz_c - hz_mx, d_hy_dx = hy_c - hy_mx;
        if (PML && live && (shell_xy || in_shell_axis(g, 2, z))) {
            if (gy) {
                const size_t i = ((size_t)z * (2 * g.T[1]) + gyi) * nx + x;
                const f4 pb = splat(esrc.prof[1][0][y]), pa = splat(esrc.prof[1][1][y]),
                         pk = splat(esrc.prof[1][2][y]);
                f4 p0 = ld4(esrc.psi[1][2] + i), p1 = ld4(esrc.psi[1][3] + i);
                d_hz_dy = cpml4(d_hz_dy, pb, pa, pk, p0);
                d_hx_dy = cpml4(d_hx_dy, pb, pa, pk, p1);
                st4(dst.psi[1][2] + i, p0);
                st4(dst.psi[1][3] + i, p1);
            } else {
                d_hz_dy = d_hz_dy + 0.0f;
                d_hx_dy = d_hx_dy + 0.0f;
            }
            if (graded(g, 2, z, g.nz)) {
                const size_t i = ((size_t)slab_index(g, 2, z, g.nz) * g.ny + y) * nx + x;
                const f4 pb = splat(esrc.prof[2][0][z]), pa = splat(esrc.prof[2][1][z]),
                         pk = splat(esrc.prof[2][2][z]);
                f4 p0 = ld4(esrc.psi[2][2] + i), p1 = ld4(esrc.psi[2][3] + i);
                d_hy_dz = cpml4(d_hy_dz, pb, pa, pk, p0);
                d_hx_dz = cpml4(d_hx_dz, pb, pa, pk, p1);
                st4(dst.psi[2][2] + i, p0);
                st4(dst.psi[2][3] + i, p1);
            } else {
                d_hy_dz = d_hy_dz + 0.0f;
                d_hx_dz = d_hx_dz + 0.0f;
            }
            if (gx) {
                const size_t i = ((size_t)z * g.ny + y) * (2 * g.T[0]) + gxi;
                const f4 pb = ld4(esrc.prof[0][0] + x), pa = ld4(esrc.prof[0][1] + x),
                         pk = ld4(esrc.prof[0][2] + x);
                f4 p0 = ld4(esrc.psi[0][2] + i), p1 = ld4(esrc.psi[0][3] + i);
                d_hz_dx = cpml4(d_hz_dx, pb, pa, pk, p0);
                d_hy_dx = cpml4(d_hy_dx, pb, pa, pk, p1);
                st4(dst.psi[0][2] + i, p0);
                st4(dst.psi[0][3] + i, p1);
            } else {
                d_hz_dx = d_hz_dx + 0.0f;
                d_hy_dx = d_hy_dx + 0.0f;
            }
        }
        if (live) {
            f4 ca, cb;
            coef4(s_co, *reinterpret_cast<const uint32_t *>(esrc.id + zo + row), ca, cb);
            const f4 cex = d_hz_dy - d_hy_dz;
            const f4 cey = d_hx_dz - d_hz_dx;
            const f4 cez = d_hy_dx - d_hx_dy;
            st4(dst.f[0] + zo + row, ca * ld4(esrc.f[0] + zo + row) + cb * cex);
            st4(dst.f[1] + zo + row, ca * ld4(esrc.f[1] + zo + row) + cb * cey);
            st4(dst.f[2] + zo + row, ca * ld4(esrc.f[2] + zo + row) + cb * cez);
        }
        hx_mz = hx_c;
        hy_mz = hy_c;
    }
}

template <bool PML, bool GHOST>
__global__ __launch_bounds__(64 * kSY) void k_e4(Geo g, Ptrs hsrc, Ptrs esrc, Ptrs dst,
                                                 BoxList bl, int zchunk)
{
    __shared__ float2 s_co[256];  // (ca, cb)
    int local;
    const Box &b = find_box(bl, xcd_tile(blockIdx.x, bl.total), local);
    e4_tile<PML, GHOST>(g, hsrc, esrc, dst, b, local, zchunk, s_co);
}

template <bool PML>
__global__ __launch_bounds__(64 * kSY) void k_e4_batch(const BatchScene *__restrict__ sc,
                                                       const uint32_t *__restrict__ wg_scene,
                                                       int total, int zchunk)
{
    __shared__ float2 s_co[256];  // (ca, cb)
    for (int i = blockIdx.x; i < total; i += gridDim.x) {
        const int t = xcd_tile(i, total);
        const BatchScene *d = sc + wg_scene[t];
        const Geo g = d->g;
        const Ptrs p = d->p;
        const Box b = d->box;
        if (t - b.first >= b.gx * b.gy * b.gz) return;  // tables and grid always agree
        e4_tile<PML>(g, p, p, p, b, t - b.first, zchunk, s_co);
        __syncthreads();  // the next tile reloads s_co
    }
}

// ===========================================================================
// Fused kernel over the interior box: H(n+1/2) and E(n+1) in one sweep,
// state A -> state B, no CPML (every cell it touches, halos included, is
// outside the PML shell, where the reference uses the plain kernels).
//
// Per cell 24 B read + 24 B written + 1 B id (the two-pass kernels move
// 74 B).  Workgroup = kFY waves x 64 lanes; wave w holds row
// y = y0 + (kFY-1)*ty - 1 + w, lane l the quad x = x0 + 256*tx + 4l.  Wave 0
// is a one-row halo: it recomputes H(n+1/2) of row y-1 from state A (the same
// arithmetic as that row's owner, so the same bits) for the E update of the
// owned rows, and never stores.  H at x-1 comes from the previous lane; lane
// 0 of every owned row also recomputes the one halo cell x0-1.  The sweep
// marches z over its slice of the box's planes, starting one plane early for
// H(z0-1); E(z+1) and H(z-1) ride in registers; H at y-1 crosses waves
// through a double-buffered LDS plane behind an LDS-only barrier, so stores
// are never waited for.
// ===========================================================================
// Waves per workgroup (owned rows + the halo row).  Plain variant: 12 -- at
// 164 VGPRs three waves fit per SIMD, so one 12-wave workgroup fills a CU
// and the halo row is 1/12 of the work and of the reads (8 waves, 1/8,
// measured 1-2% slower per launch at 1024^3 on the same allocation).  The
// CPML and ghost variants need more registers and keep 8.
#ifndef WF_FUSED_FY
#define WF_FUSED_FY 12
#endif
constexpr int kFY = WF_FUSED_FY;
constexpr int kFYx = 8;
// PM: 0 no CPML cells, 1 CPML cells on any axis, 2 CPML cells on the x
// axis only (a box whose y and z ranges lie outside the shell).
#ifndef WF_FUSED_XPML_FY
#define WF_FUSED_XPML_FY 8  // 198 VGPRs: 12 waves would spill
#endif
// The ghost variant (learned faces) at 8 waves, 182 VGPRs: at 12 waves it
// spills 14 and the learned 256^3 step took 0.281 against 0.267 ms.
#ifndef WF_FUSED_GHOST_FY
#define WF_FUSED_GHOST_FY 8
#endif
template <int PM, bool GHOST>
constexpr int fused_waves()
{
    return PM == 1 ? kFYx : (GHOST ? WF_FUSED_GHOST_FY : (PM == 2 ? WF_FUSED_XPML_FY : kFY));
}

// CPML of one derivative of a cell (quad) in the PML shell on axis `ax`:
// graded -> the one-tap recursion with psi read from `pin` and, when
// `store`, the new psi written to `pout`; otherwise raw + 0.0f, the
// reference's identity profile (src/fdtd.rs:136-163).
__device__ __forceinline__ f4 pml_d4(f4 raw, bool gr, f4 pb, f4 pa, f4 pk, const float *pin,
                                     float *pout, size_t i, bool store)
{
    if (!gr) return raw + 0.0f;
    f4 p = ld4(pin + i);
    const f4 d = cpml4(raw, pb, pa, pk, p);
    if (store) st4(pout + i, p);
    return d;
}

__device__ __forceinline__ float pml_d1(float raw, bool gr, float pb, float pa, float pk,
                                        const float *pin, size_t i)
{
    if (!gr) return raw + 0.0f;
    float p = pin[i];
    return cpml(raw, pb, pa, pk, p);
}

// GHOST: the E planes just beyond the high faces come from learned ghost
// planes where present (a.ghost[1], [3], [5]: E that the H phase of the
// boundary cells reads); the low faces' ghost H is not known yet when this
// kernel runs, so their boundary cells' E is redone afterwards by the E
// kernel (see the engine's learned-face schedule).
template <int PM, bool GHOST>
__global__ __launch_bounds__((64 * fused_waves<PM, GHOST>())) void k_fused4(Geo g, Ptrs a, Ptrs b,
                                                                        Box box)
{
    constexpr bool PML = PM != 0;   // the box may hold CPML shell cells
    constexpr bool PYZ = PM == 1;   // ... on the y and z axes too
    constexpr int FY = fused_waves<PM, GHOST>();
    __shared__ float2 s_ce[256];      // (ca, cb)
    __shared__ float2 s_ch[256];      // (da, db)
    __shared__ f4 s_h[2][2][FY][64];  // [buf][hz, hx][wave][lane]
    const int lane = threadIdx.x, w = threadIdx.y;
    for (int i = w * 64 + lane; i < 256; i += 64 * FY) {
        const float4 c = a.coef[i];
        s_ce[i] = make_float2(c.x, c.y);
        s_ch[i] = make_float2(c.z, c.w);
    }
    __syncthreads();

    const int t = xcd_tile(blockIdx.x, box.gx * box.gy * box.gz);
    const int tx = t % box.gx, ty = (t / box.gx) % box.gy, tz = t / (box.gx * box.gy);
    const int x = box.x0 + tx * box.tstride + lane * 4;
    const int yr = box.y0 + ty * (FY - 1) - 1 + w;
    int y = yr;
    bool live = x < box.x1 && yr < box.y1;
    if (yr < 0) {
        if (g.periodic & 2u) y = g.ny - 1; else live = false;
    }
    const bool own = live && w >= 1;
    const int xc = live ? x : 0, yc = live ? y : 0;
    const size_t nx = g.nx, plane = nx * g.ny;
    const size_t row = (size_t)yc * nx + xc;
    const int yp = plus1(yc, g.ny, g.periodic & 2u);
    const size_t rowp_y = (size_t)(yp >= 0 ? yp : 0) * nx + xc;
    const bool edge = lane == 63 || x + 4 >= box.x1;
    const int xq = plus1(xc + 3, g.nx, g.periodic & 1u);
    const size_t rowq = (size_t)yc * nx + (xq >= 0 ? xq : 0);
    // x-halo cell x0-1 for lane 0 of the owned rows
    const int xh = minus1(xc, g.nx, g.periodic & 1u);
    const bool hlive = own && lane == 0 && xh >= 0;
    const size_t rowh = (size_t)yc * nx + (xh >= 0 ? xh : 0);
    const size_t rowh_py = (size_t)(yp >= 0 ? yp : 0) * nx + (xh >= 0 ? xh : 0);

    // CPML shell and grading of this lane's quad (x), this wave's row (y),
    // and the halo cell; profiles of graded positions (identity elsewhere).
    bool sh_x = false, sh_y = false, sh_h = false, gx = false, gy = false, gh = false;
    int gxi = 0, gyi = 0, ghi = 0;
    f4 pbx = splat(1.f), pax = splat(0.f), pkx = splat(1.f);
    float pby = 1.f, pay = 0.f, pky = 1.f, pbh = 1.f, pah = 0.f, pkh = 1.f;
    if (PML && live) {
        sh_x = in_shell_axis(g, 0, xc);
        gx = graded(g, 0, xc, g.nx);
        gxi = slab_index(g, 0, xc, g.nx);
        if (PYZ) {
            sh_y = in_shell_axis(g, 1, yc);
            gy = graded(g, 1, yc, g.ny);
            gyi = slab_index(g, 1, yc, g.ny);
        }
        if (gx && PYZ) {  // x-only boxes load them where they are used
            pbx = ld4(a.prof[0][0] + xc);
            pax = ld4(a.prof[0][1] + xc);
            pkx = ld4(a.prof[0][2] + xc);
        }
        if (gy) {
            pby = a.prof[1][0][yc];
            pay = a.prof[1][1][yc];
            pky = a.prof[1][2][yc];
        }
        if (hlive && PYZ) {  // x-only boxes keep their x-halo cell outside the shell
            sh_h = in_shell_axis(g, 0, xh);
            gh = graded(g, 0, xh, g.nx);
            ghi = slab_index(g, 0, xh, g.nx);
            if (gh) {
                pbh = a.prof[0][0][xh];
                pah = a.prof[0][1][xh];
                pkh = a.prof[0][2][xh];
            }
        }
    }
    const size_t Tx2 = 2 * (size_t)g.T[0], Ty2 = 2 * (size_t)g.T[1];

    const float *__restrict__ aex = a.f[0];
    const float *__restrict__ aey = a.f[1];
    const float *__restrict__ aez = a.f[2];
    const float *__restrict__ ahx = a.f[3];
    const float *__restrict__ ahy = a.f[4];
    const float *__restrict__ ahz = a.f[5];

    // the box's planes split evenly into gz chunks (sizes differ by <= 1)
    const long nzb = box.z1 - box.z0;
    const int z0 = box.z0 + (int)(tz * nzb / box.gz), z1 = box.z0 + (int)((tz + 1) * nzb / box.gz);
    const int zpre = minus1(z0, g.nz, g.periodic & 4u);
    const int it0 = zpre >= 0 ? -1 : 0;
    const int nit = z1 - z0;

    // Everything one plane z contributes from state A: E at z+1 and at y+1,
    // H at z, the material ids, the edge lane's x+1 cells, and the x-halo
    // cell's inputs.  The next plane's copy is loaded before the current
    // plane is computed, so its latency overlaps the arithmetic, the LDS
    // exchange and the stores.
    struct Plane {
        f4 ex_n, ey_n, ez_n, ez_py, ex_py, hx_o, hy_o, hz_o;
        uint32_t ids;
        float sez, sey;                           // edge lane: E at x+4
        float hexn, heyn, hezn, hexpy, hhy, hhz;  // x-halo cell
        int hid;
        f4 psx[PM == 2 ? 4 : 1];                  // x-only boxes: the quad's x psi terms
    };
    auto plane_z = [&](int it) { return it < 0 ? zpre : z0 + it; };
    auto load = [&](Plane &P, int it) {
        const int z = plane_z(it);
        const size_t zo = (size_t)z * plane;
        const int zp = plus1(z, g.nz, g.periodic & 4u);
        P.ex_n = P.ey_n = P.ez_n = P.ez_py = P.ex_py = P.hx_o = P.hy_o = P.hz_o = splat(0.f);
        P.ids = 0;
        P.sez = P.sey = 0.f;
        P.hexn = P.heyn = P.hezn = P.hexpy = P.hhy = P.hhz = 0.f;
        P.hid = 0;
        if (PM == 2) {
            // the x psi terms of a graded quad, loaded a plane ahead like the
            // fields (a quad is graded or not as a whole: T is whole blocks)
            for (int t = 0; t < (PM == 2 ? 4 : 1); ++t) P.psx[t] = splat(0.f);
            if (live && gx) {
                const size_t jx = ((size_t)plane_z(it) * g.ny + yc) * (2 * (size_t)g.T[0]) + gxi;
                for (int t = 0; t < (PM == 2 ? 4 : 1); ++t) P.psx[t] = ld4(a.psi[0][t] + jx);
            }
        }
        if (live) {
            if (zp >= 0) {
                const size_t on = (size_t)zp * plane + row;
                P.ex_n = ld4(aex + on);
                P.ey_n = ld4(aey + on);
                P.ez_n = ld4(aez + on);
            } else if (GHOST && a.ghost[5][0]) {
                P.ex_n = ld4(a.ghost[5][0] + (size_t)yc * nx + xc);
                P.ey_n = ld4(a.ghost[5][1] + (size_t)yc * nx + xc);
            }
            if (yp >= 0) {
                P.ez_py = ld4(aez + zo + rowp_y);
                P.ex_py = ld4(aex + zo + rowp_y);
            } else if (GHOST && a.ghost[3][0]) {
                P.ez_py = ld4(a.ghost[3][0] + (size_t)z * nx + xc);
                P.ex_py = ld4(a.ghost[3][1] + (size_t)z * nx + xc);
            }
            P.hx_o = ld4(ahx + zo + row);
            if (w >= 1) P.hy_o = ld4(ahy + zo + row);  // the halo row needs hx, hz only
            P.hz_o = ld4(ahz + zo + row);
            P.ids = *reinterpret_cast<const uint32_t *>(a.id + zo + row);
            if (edge && xq >= 0) {
                P.sez = aez[zo + rowq];
                P.sey = aey[zo + rowq];
            } else if (GHOST && edge && a.ghost[1][0]) {
                P.sey = a.ghost[1][0][(size_t)z * g.ny + yc];
                P.sez = a.ghost[1][1][(size_t)z * g.ny + yc];
            }
        }
        if (hlive) {
            if (zp >= 0) {
                const size_t on = (size_t)zp * plane + rowh;
                P.hexn = aex[on];
                P.heyn = aey[on];
                P.hezn = aez[on];
            } else if (GHOST && a.ghost[5][0]) {
                P.hexn = a.ghost[5][0][(size_t)yc * nx + xh];
                P.heyn = a.ghost[5][1][(size_t)yc * nx + xh];
            }
            if (it >= 0) {
                P.hexpy = yp >= 0 ? aex[zo + rowh_py]
                                  : ((GHOST && a.ghost[3][1]) ? a.ghost[3][1][(size_t)z * nx + xh] : 0.f);
                P.hid = a.id[zo + rowh];
                P.hhy = ahy[zo + rowh];
                P.hhz = ahz[zo + rowh];
            }
        }
    };

    // E^n of the own quad (and of the halo cell) at the current plane;
    // H(n+1/2) of the own quad one plane below.
    f4 ex_c = splat(0.f), ey_c = splat(0.f), ez_c = splat(0.f);
    f4 hx_mz = splat(0.f), hy_mz = splat(0.f);
    float hex = 0.f, hey = 0.f, hez = 0.f;
    {
        const size_t o = (size_t)plane_z(it0) * plane;
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
    Plane cur;
    load(cur, it0);

    for (int it = it0; it < nit; ++it) {
        const int z = plane_z(it);
        const size_t zo = (size_t)z * plane;
        Plane nxt;
        if (it + 1 < nit) load(nxt, it + 1);

        // CPML of this plane (uniform): shell on z, grading, z profile
        bool sh_z = false, gz = false, pml_q = false;
        size_t gzo = 0;
        float pbz = 1.f, paz = 0.f, pkz = 1.f;
        if (PML) {
            if (PYZ) {
                sh_z = in_shell_axis(g, 2, z);
                gz = graded(g, 2, z, g.nz);
            }
            if (gz) {
                gzo = (size_t)slab_index(g, 2, z, g.nz) * plane;
                pbz = a.prof[2][0][z];
                paz = a.prof[2][1][z];
                pkz = a.prof[2][2][z];
            }
            pml_q = live && (sh_x || sh_y || sh_z);
        }
        const bool store = own && it >= 0;  // the pre-plane and halo row only read psi
        const size_t iy = ((size_t)z * Ty2 + gyi) * nx + xc;  // psi indices of the quad
        const size_t iz = gzo + row;
        const size_t ix = ((size_t)z * g.ny + yc) * Tx2 + gxi;

        float sez = __shfl_down(ez_c.x, 1, 64), sey = __shfl_down(ey_c.x, 1, 64);
        if (edge) {
            sez = cur.sez;
            sey = cur.sey;
        }

        // ---- H(n+1/2) of the own quad: update_h_field[_pml] ------------------
        const f4 ez_px = f4{ez_c.y, ez_c.z, ez_c.w, sez};
        const f4 ey_px = f4{ey_c.y, ey_c.z, ey_c.w, sey};
        f4 da, db;
        coef4(s_ch, cur.ids, da, db);
        f4 d_ez_dy = cur.ez_py - ez_c, d_ey_dz = cur.ey_n - ey_c;
        f4 d_ex_dz = cur.ex_n - ex_c, d_ez_dx = ez_px - ez_c;
        f4 d_ey_dx = ey_px - ey_c, d_ex_dy = cur.ex_py - ex_c;
        if (PML && pml_q) {
            d_ez_dy = pml_d4(d_ez_dy, gy, splat(pby), splat(pay), splat(pky), a.psi[1][0],
                             b.psi[1][0], iy, store);
            d_ex_dy = pml_d4(d_ex_dy, gy, splat(pby), splat(pay), splat(pky), a.psi[1][1],
                             b.psi[1][1], iy, store);
            d_ey_dz = pml_d4(d_ey_dz, gz, splat(pbz), splat(paz), splat(pkz), a.psi[2][0],
                             b.psi[2][0], iz, store);
            d_ex_dz = pml_d4(d_ex_dz, gz, splat(pbz), splat(paz), splat(pkz), a.psi[2][1],
                             b.psi[2][1], iz, store);
            if (PYZ) {
                d_ez_dx = pml_d4(d_ez_dx, gx, pbx, pax, pkx, a.psi[0][0], b.psi[0][0], ix, store);
                d_ey_dx = pml_d4(d_ey_dx, gx, pbx, pax, pkx, a.psi[0][1], b.psi[0][1], ix, store);
            } else if (gx) {
                const f4 qb = ld4(a.prof[0][0] + xc), qa = ld4(a.prof[0][1] + xc),
                         qk = ld4(a.prof[0][2] + xc);
                f4 p0 = cur.psx[0], p1 = cur.psx[PM == 2 ? 1 : 0];
                d_ez_dx = cpml4(d_ez_dx, qb, qa, qk, p0);
                d_ey_dx = cpml4(d_ey_dx, qb, qa, qk, p1);
                if (store) {
                    st4(b.psi[0][0] + ix, p0);
                    st4(b.psi[0][1] + ix, p1);
                }
            } else {
                d_ez_dx = d_ez_dx + 0.0f;
                d_ey_dx = d_ey_dx + 0.0f;
            }
        }
        f4 hxn = da * cur.hx_o - db * (d_ez_dy - d_ey_dz);
        f4 hyn = da * cur.hy_o - db * (d_ex_dz - d_ez_dx);
        f4 hzn = da * cur.hz_o - db * (d_ey_dx - d_ex_dy);
        if (!live) {
            hxn = splat(0.f);
            hyn = splat(0.f);
            hzn = splat(0.f);
        }

        if (it >= 0) {
            // ---- H(n+1/2) of the x-halo cell (hy, hz only) ------------------
            float hy_h = 0.f, hz_h = 0.f;
            if (hlive) {
                const float2 c = s_ch[cur.hid];
                float h_ex_dz = cur.hexn - hex, h_ez_dx = ez_c.x - hez;
                float h_ey_dx = ey_c.x - hey, h_ex_dy = cur.hexpy - hex;
                if (PYZ && (sh_h || sh_y || sh_z)) {
                    const size_t jx = ((size_t)z * g.ny + yc) * Tx2 + ghi;
                    const size_t jy = ((size_t)z * Ty2 + gyi) * nx + xh;
                    h_ex_dz = pml_d1(h_ex_dz, gz, pbz, paz, pkz, a.psi[2][1], gzo + rowh);
                    h_ez_dx = pml_d1(h_ez_dx, gh, pbh, pah, pkh, a.psi[0][0], jx);
                    h_ey_dx = pml_d1(h_ey_dx, gh, pbh, pah, pkh, a.psi[0][1], jx);
                    h_ex_dy = pml_d1(h_ex_dy, gy, pby, pay, pky, a.psi[1][1], jy);
                }
                hy_h = c.x * cur.hhy - c.y * (h_ex_dz - h_ez_dx);
                hz_h = c.x * cur.hhz - c.y * (h_ey_dx - h_ex_dy);
            }
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
            if (own) {
                const f4 hz_my = s_h[buf][0][w - 1][lane];
                const f4 hx_my = s_h[buf][1][w - 1][lane];
                const f4 hz_mx = f4{hz_l, hzn.x, hzn.y, hzn.z};
                const f4 hy_mx = f4{hy_l, hyn.x, hyn.y, hyn.z};
                // ---- E(n+1): update_e_field[_pml] ---------------------------
                f4 ca, cb;
                coef4(s_ce, cur.ids, ca, cb);
                f4 d_hz_dy = hzn - hz_my, d_hy_dz = hyn - hy_mz;
                f4 d_hx_dz = hxn - hx_mz, d_hz_dx = hzn - hz_mx;
                f4 d_hy_dx = hyn - hy_mx, d_hx_dy = hxn - hx_my;
                if (PML && pml_q) {
                    d_hz_dy = pml_d4(d_hz_dy, gy, splat(pby), splat(pay), splat(pky),
                                     a.psi[1][2], b.psi[1][2], iy, true);
                    d_hx_dy = pml_d4(d_hx_dy, gy, splat(pby), splat(pay), splat(pky),
                                     a.psi[1][3], b.psi[1][3], iy, true);
                    d_hy_dz = pml_d4(d_hy_dz, gz, splat(pbz), splat(paz), splat(pkz),
                                     a.psi[2][2], b.psi[2][2], iz, true);
                    d_hx_dz = pml_d4(d_hx_dz, gz, splat(pbz), splat(paz), splat(pkz),
                                     a.psi[2][3], b.psi[2][3], iz, true);
                    if (PYZ) {
                        d_hz_dx = pml_d4(d_hz_dx, gx, pbx, pax, pkx, a.psi[0][2], b.psi[0][2], ix,
                                         true);
                        d_hy_dx = pml_d4(d_hy_dx, gx, pbx, pax, pkx, a.psi[0][3], b.psi[0][3], ix,
                                         true);
                    } else if (gx) {
                        const f4 qb = ld4(a.prof[0][0] + xc), qa = ld4(a.prof[0][1] + xc),
                                 qk = ld4(a.prof[0][2] + xc);
                        f4 p2 = cur.psx[PM == 2 ? 2 : 0], p3 = cur.psx[PM == 2 ? 3 : 0];
                        d_hz_dx = cpml4(d_hz_dx, qb, qa, qk, p2);
                        d_hy_dx = cpml4(d_hy_dx, qb, qa, qk, p3);
                        st4(b.psi[0][2] + ix, p2);
                        st4(b.psi[0][3] + ix, p3);
                    } else {
                        d_hz_dx = d_hz_dx + 0.0f;
                        d_hy_dx = d_hy_dx + 0.0f;
                    }
                }
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
        ex_c = cur.ex_n;
        ey_c = cur.ey_n;
        ez_c = cur.ez_n;
        hex = cur.hexn;
        hey = cur.heyn;
        hez = cur.hezn;
        cur = nxt;
    }
}

// ===========================================================================
// Two-step kernel: two whole timesteps of a box in one sweep, state A (E^n,
// H^{n-1/2}) -> state B (E^{n+2}, H^{n+3/2}).  Per owned cell 24 B read +
// 24 B written + 1 B id (+ 8 B per psi term) for the two steps: half the
// single-pass kernel's bytes per step.
//
// Rows of LPR lanes (64 / LPR rows per wave); lane l of a row holds the quad
// x = tx0 - 4 + 4l, row R of the tile (R = wave * rows per wave + row) the
// row y = y0 + (rows - 3) ty - 2 + R.  Two steps reach two cells down and one
// up on each axis, so a tile also computes a margin that it never stores:
// the first and last lane of each row (a quad on each side), rows 0 and 1
// (two below), the last row (one above), two planes below its z slice and
// one above.  The outermost margin values are wrong (their neighbours lie
// outside the tile), and each step reads them only for margin cells one
// cell further out:
//   step 1 at plane p: H(n+1/2) on every row, E(n+1) on rows >= 1;
//   step 2 at plane p-1 (H(n+3/2) at p-1 reads E(n+1) at p): H(n+3/2) on
//   rows 1..last-1, E(n+2) on rows 2..last-1, the rows the tile owns.
// y-neighbours cross rows through LDS (H(n+1/2) at y-1, E(n+1) at y+1 one
// plane back, H(n+3/2) at y-1, E^n of the next plane at y+1; two LDS-only
// barriers per plane), x-neighbours come from the adjacent lane,
// z-neighbours ride in registers.
// The sources of step n+1 are added to E(n+1) wherever it is computed (rows
// flagged in a bitmap, entries in list order), and the probes of step n+1
// at owned cells are recorded into their slots.  The arithmetic of every
// update is the single-pass kernel's, so the bits are the same.
// ===========================================================================
#ifndef WF_TWOSTEP_FY
#define WF_TWOSTEP_FY 15
#endif
constexpr int kF2 = WF_TWOSTEP_FY;

// First entry of row `key` in a point table (entries sorted by row).
__device__ __forceinline__ int point_row_begin(const PointRows &t, uint64_t key)
{
    int lo = 0, hi = t.n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (t.key[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__device__ __forceinline__ bool point_row_flagged(const PointRows &t, uint64_t key)
{
    return t.bits && ((t.bits[key >> 5] >> (key & 31)) & 1u);
}

// a*b + c for the field updates of the two-step kernel, a the field's own
// coefficient and b its old value.  The contracted set fuses exactly this
// product (the coefficient is 1 in lossless cells, where the fused result
// then equals the separately rounded one); the exact set rounds the product
// and the sum separately, and a*b + (-(d*e)) is a*b - d*e bit for bit.
__device__ __forceinline__ f4 madd4(f4 a, f4 b, f4 c)
{
#ifdef WF_STEP_FMA
    return __builtin_elementwise_fma(a, b, c);
#else
    return a * b + c;
#endif
}

// Element k (0..3) of a quad, read or replaced without dynamic indexing.
__device__ __forceinline__ float quad_get(const f4 &v, int k)
{
    return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w));
}

// src/source.rs:79-95: the field value += the waveform value, or = it.
__device__ __forceinline__ void quad_source(f4 &v, int k, float val, bool hard)
{
    const float nv = hard ? val : quad_get(v, k) + val;
    v.x = k == 0 ? nv : v.x;
    v.y = k == 1 ? nv : v.y;
    v.z = k == 2 ? nv : v.z;
    v.w = k == 3 ? nv : v.w;
}

// PM: the CPML axes of the cells a variant computes, bit 1 x, 2 y, 4 z (the
// domain is cut into regions, shell or not on each axis, and each region
// runs the variant of its axes); LPRT: lanes per row, fixed (64, 32) or box.lpr
// (0: the x CPML regions, whose rows are narrow), several rows per wave when
// narrow.  CPML variants carry each psi term of a graded cell from the first
// step to the second in registers; every variant holds lanes, rows and
// planes outside the domain at zero, the boundary the reference reads there.
constexpr int kF2x = 8;  // waves of the CPML variants (their psi terms need the registers)
// The interior variant loads each plane when it reaches it, with 15 waves
// per workgroup (127 VGPRs: four waves per SIMD), reads the previous plane's
// E(n+1) back from its own LDS slot, and takes E^n of the row above from the
// row above through LDS (that row loaded it as its own next plane).  Measured
// on MI355X at 1024^3 (A/B on one GPU): one plane loaded ahead with 12
// waves (167 VGPRs, three per SIMD) ran 7.85-7.91 ms per step against
// 7.53-7.65 for this scheme at 16 waves; reading back also H(n+1/2) and
// H(n+3/2) gained nothing; two planes ahead spill; loading E^n at y+1 again
// fetched it from HBM a second time (31.4 against 27.4 GB read per launch:
// the row above loaded it a plane earlier, and a plane of an XCD's 32 tiles
// fills its L2).  15 waves keep the LDS within 160 KB.
// The x CPML variant <1> (the x shell columns) can carry its x psi(n+1) in
// LDS instead of registers, which lets it run more waves: WF_TWOSTEP_FYX
// waves, the carry on when that is above kF2x.  At 10 waves it holds 168
// VGPRs with 10 spilled, and at 1024^3 its launches took 1.81-1.83 ms against
// 1.71-1.73 for registers at 8 waves (one GPU, alternating runs), which stay
// the default.
#ifndef WF_TWOSTEP_FYX
#define WF_TWOSTEP_FYX 8
#endif
template <int PM>
constexpr int twostep_waves() { return PM == 0 ? kF2 : (PM == 1 ? WF_TWOSTEP_FYX : kF2x); }
template <int PM>
constexpr int twostep_carry() { return PM == 1 && WF_TWOSTEP_FYX > kF2x ? 1 : 0; }

// One tile of the two-step sweep with FY waves: tile t of `box` (x fastest,
// then y, then the z slice), state a -> b.  QA (a CPML axis of the variant,
// bit 1 x, 2 y, 4 z; 0 none): that axis's psi(n+1), carried to the second
// step, waits in the lane's own LDS slot, each half swapped in where the
// second step uses it, instead of in registers.
template <int PM, int LPRT, int FY, int QA = 0>
__device__ __forceinline__ void twostep_tile(const Geo &g, const Ptrs &a, const Ptrs &b,
                                             const Box &box, const TwoStepPoints &pts, int t)
{
    constexpr bool PX = (PM & 1) != 0, PY = (PM & 2) != 0, PZ = (PM & 4) != 0;
    constexpr int NX = PX ? 4 : 1, NY = PY ? 4 : 1, NZ = PZ ? 4 : 1;  // psi terms per quad
    __shared__ float2 s_ce[256];        // (ca, cb)
    __shared__ float2 s_ch[256];        // (da, db)
    __shared__ f4 s_h1[2][FY * 64];     // H(n+1/2) of plane p: [hz, hx][slot]
    __shared__ f4 s_e1[2][2][FY * 64];  // E(n+1): [plane parity][ez, ex][slot]
    __shared__ f4 s_h2[2][FY * 64];     // H(n+3/2) of plane p-1: [hz, hx][slot]
    __shared__ f4 s_ep[2][FY * 64];     // E^n of the next plane: [ez, ex][slot]
    constexpr bool QL = QA == 1 && PX, QY = QA == 2 && PY, QZ = QA == 4 && PZ;
    __shared__ f4 s_q[QA ? 4 : 1][QA ? FY * 64 : 1];  // psi(n+1) of axis QA at p-1, own slot
    const int lane = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.y);
    for (int i = w * 64 + lane; i < 256; i += 64 * FY) {
        const float4 c = a.coef[i];
        s_ce[i] = make_float2(c.x, c.y);
        s_ch[i] = make_float2(c.z, c.w);
    }

    // lanes per row, rows per wave; row R of the tile = w * rpw + r (lanes
    // past rpw * LPR idle); LDS slot of each lane and of its y-1 / y+1 rows
    const int LPR = LPRT ? LPRT : box.lpr;
    const int rpw = 64 / LPR;
    const int r = LPRT == 64 ? 0 : lane / LPR, l = lane - r * LPR;
    const bool idle = r >= rpw;
    const int RT = FY * rpw;
    const int R = idle ? -8 : w * rpw + r;
    const int me = w * 64 + lane;
    const int below = (idle || R == 0) ? me : (r > 0 ? me - LPR : (w - 1) * 64 + (rpw - 1) * LPR + l);
    const int above = (idle || R == RT - 1) ? me : (r < rpw - 1 ? me + LPR : (w + 1) * 64 + l);

    const int tx = t % box.gx, ty = (t / box.gx) % box.gy, tz = t / (box.gx * box.gy);
    const int tx0 = box.x0 + tx * box.tstride;             // the tile's owned cells: [tx0, tx1)
    const int tx1 = min(tx0 + 4 * (LPR - 2), box.x1);
    const int x = tx0 - 4 + 4 * l;
    const int y = box.y0 + (RT - 3) * ty - 2 + R;
    // tx1 and y1 themselves hold the upper margin; beyond them, outside the
    // domain and on idle lanes nothing is loaded and every value is zero
    const bool live_row = !idle && y >= 0 && y < g.ny && y <= box.y1;
    const bool live = live_row && x >= 0 && x + 4 <= g.nx && x <= tx1;
    const bool own_row = live_row && y < box.y1 && R >= 2 && R <= RT - 2;
    const bool own = own_row && live && x >= tx0 && x < tx1;
    const size_t nx = g.nx, plane = nx * g.ny;
    const size_t row = (size_t)(live ? y : 0) * nx + (live ? x : 0);
    const size_t rowp = row + nx;  // E^n at y+1
    // E^n at y+1: inside the domain, and loaded by the rows whose row above
    // is not in the tile (the others take it from that row through LDS)
    const bool has_py = y + 1 < g.ny;
    const bool epy_mem = R == RT - 1 || y >= box.y1;

    const long nzb = box.z1 - box.z0;
    const int z0 = box.z0 + (int)(tz * nzb / box.gz), z1 = box.z0 + (int)((tz + 1) * nzb / box.gz);
    const int nit = z1 - z0;

    // step n+1's row of source values
    const float *sv = nullptr;
    if (pts.src.bits) sv = pts.values + (*pts.step - pts.base) * (uint64_t)pts.ncol;

    // CPML of this lane's quad (x) and this row (y): in the shell (identity
    // derivatives + 0.0f) and graded (the one-tap recursion; the shell is the
    // graded layer, a quad lies in one block); psi offsets and profiles
    bool shx = false, gxq = false, shy = false, gyr = false;
    size_t gxi = 0, gyo = 0;
    f4 pbx = splat(1.f), pax = splat(0.f), pkx = splat(1.f);
    float pby = 1.f, pay = 0.f, pky = 1.f;
    if (PX && live) {
        shx = in_shell_axis(g, 0, x);
        gxq = graded(g, 0, x, g.nx);
        if (gxq) {
            gxi = (size_t)y * 2 * (size_t)g.T[0] + (size_t)slab_index(g, 0, x, g.nx);
            pbx = ld4(a.prof[0][0] + x);
            pax = ld4(a.prof[0][1] + x);
            pkx = ld4(a.prof[0][2] + x);
        }
    }
    if (PY && live) {
        shy = in_shell_axis(g, 1, y);
        gyr = graded(g, 1, y, g.ny);
        if (gyr) {
            gyo = (size_t)slab_index(g, 1, y, g.ny) * nx + x;
            pby = a.prof[1][0][y];
            pay = a.prof[1][1][y];
            pky = a.prof[1][2][y];
        }
    }
    const size_t psx = (size_t)g.ny * 2 * (size_t)g.T[0];  // x psi per plane
    const size_t psy = 2 * (size_t)g.T[1] * nx;             // y psi per plane

    const float *__restrict__ aex = a.f[0];
    const float *__restrict__ aey = a.f[1];
    const float *__restrict__ aez = a.f[2];
    const float *__restrict__ ahx = a.f[3];
    const float *__restrict__ ahy = a.f[4];
    const float *__restrict__ ahz = a.f[5];

    // State A at plane p: E at p+1 and (rows whose row above is not in the
    // tile) at y+1, H, the material ids, the psi terms of graded cells.
    struct Plane {
        f4 ex_n, ey_n, ez_n, ez_py, ex_py, hx_o, hy_o, hz_o;
        uint32_t ids;
        f4 px[NX], py[NY], pz[NZ];
    };
    auto load = [&](Plane &P, int p) {
        P.ex_n = P.ey_n = P.ez_n = P.ez_py = P.ex_py = P.hx_o = P.hy_o = P.hz_o = splat(0.f);
        P.ids = 0;
        for (int k = 0; k < NX; ++k) P.px[k] = splat(0.f);
        for (int k = 0; k < NY; ++k) P.py[k] = splat(0.f);
        for (int k = 0; k < NZ; ++k) P.pz[k] = splat(0.f);
        // E at p+1 becomes the next plane's E^n: loaded whenever p+1 is in
        // the domain, also for a plane p below it
        if (live && p + 1 >= 0 && p + 1 < g.nz) {
            const size_t on = (size_t)(p + 1) * plane + row;
            P.ex_n = ld4(aex + on);
            P.ey_n = ld4(aey + on);
            P.ez_n = ld4(aez + on);
        }
        if (live && p >= 0 && p < g.nz) {
            const size_t zo = (size_t)p * plane;
            if (epy_mem && has_py) {
                P.ez_py = ld4(aez + zo + rowp);
                P.ex_py = ld4(aex + zo + rowp);
            }
            P.hx_o = ld4(ahx + zo + row);
            if (R >= 1) P.hy_o = ld4(ahy + zo + row);  // row y0-2 needs H(n+1/2) hz, hx only
            P.hz_o = ld4(ahz + zo + row);
            P.ids = *reinterpret_cast<const uint32_t *>(a.id + zo + row);
            if (PX && gxq)
                for (int k = 0; k < NX; ++k) P.px[k] = ld4(a.psi[0][k] + (size_t)p * psx + gxi);
            if (PY && gyr)
                for (int k = 0; k < NY; ++k) P.py[k] = ld4(a.psi[1][k] + (size_t)p * psy + gyo);
            if (PZ && graded(g, 2, p, g.nz)) {
                const size_t o = (size_t)slab_index(g, 2, p, g.nz) * plane + row;
                for (int k = 0; k < NZ; ++k) P.pz[k] = ld4(a.psi[2][k] + o);
            }
        }
    };

    f4 ex_c = splat(0.f), ey_c = splat(0.f), ez_c = splat(0.f);  // E^n at p
    if (live && z0 - 2 >= 0) {
        const size_t o = (size_t)(z0 - 2) * plane + row;
        ex_c = ld4(aex + o);
        ey_c = ld4(aey + o);
        ez_c = ld4(aez + o);
    }
    s_ep[0][me] = ez_c;
    s_ep[1][me] = ex_c;
    __syncthreads();
    f4 h1x_c = splat(0.f), h1y_c = splat(0.f), h1z_c = splat(0.f);  // H(n+1/2) at p-1
    f4 e1x_c, e1y_c = splat(0.f), e1z_c;                           // E(n+1) at p-1
    f4 h2x_m = splat(0.f), h2y_m = splat(0.f);                      // H(n+3/2) at p-2
    f4 qx[NX], qy[NY], qz[NZ];  // psi(n+1) at p-1
    for (int k = 0; k < NX; ++k) qx[k] = splat(0.f);
    if (QA)
        for (int k = 0; k < 4; ++k) s_q[k][me] = splat(0.f);
    for (int k = 0; k < NY; ++k) qy[k] = splat(0.f);
    for (int k = 0; k < NZ; ++k) qz[k] = splat(0.f);
    uint32_t ids_c = 0;                       // ids at p-1
    bool shz_c = false, gz_c = false;         // z CPML of plane p-1
    float pbz_c = 1.f, paz_c = 0.f, pkz_c = 1.f;
    Plane cur;

    for (int it = -2; it <= nit; ++it) {
        const int p = z0 + it;
        const bool lp = p >= 0 && p < g.nz;      // plane p in the domain
        const bool lq = p >= 1 && p <= g.nz;     // plane p-1 in the domain
        load(cur, p);
        // z CPML of plane p
        bool shz = false, gzp = false;
        float pbz = 1.f, paz = 0.f, pkz = 1.f;
        if (PZ && lp) {
            shz = in_shell_axis(g, 2, p);
            gzp = graded(g, 2, p, g.nz);
            if (gzp) {
                pbz = a.prof[2][0][p];
                paz = a.prof[2][1][p];
                pkz = a.prof[2][2][p];
            }
        }
        const bool sh1 = shx || shy || shz;   // the cell's block is in the shell (plane p)
        const bool sh2 = shx || shy || shz_c; // ... at plane p-1
        f4 q1x[NX], q1y[NY], q1z[NZ];  // psi(n+1) at p
        for (int k = 0; k < NX; ++k) q1x[k] = cur.px[k];
        for (int k = 0; k < NY; ++k) q1y[k] = cur.py[k];
        for (int k = 0; k < NZ; ++k) q1z[k] = cur.pz[k];

        // ---- step 1: H(n+1/2) at plane p (update_h_field[_pml]) ---------
        const float sez = __shfl_down(ez_c.x, 1, 64), sey = __shfl_down(ey_c.x, 1, 64);
        const f4 ez_px = f4{ez_c.y, ez_c.z, ez_c.w, sez};
        const f4 ey_px = f4{ey_c.y, ey_c.z, ey_c.w, sey};
        f4 h1x, h1y, h1z;
        {
            f4 da, db;
            coef4(s_ch, cur.ids, da, db);
            f4 ez_py = cur.ez_py, ex_py = cur.ex_py;
            if (!epy_mem) {
                ez_py = s_ep[0][above];
                ex_py = s_ep[1][above];
            }
            f4 d_ez_dy = ez_py - ez_c, d_ey_dz = cur.ey_n - ey_c;
            f4 d_ex_dz = cur.ex_n - ex_c, d_ez_dx = ez_px - ez_c;
            f4 d_ey_dx = ey_px - ey_c, d_ex_dy = ex_py - ex_c;
            if (PM && sh1) {
                if (PY && gyr) {
                    d_ez_dy = cpml4(d_ez_dy, splat(pby), splat(pay), splat(pky), q1y[0]);
                    d_ex_dy = cpml4(d_ex_dy, splat(pby), splat(pay), splat(pky), q1y[PY ? 1 : 0]);
                } else {
                    d_ez_dy = d_ez_dy + 0.0f;
                    d_ex_dy = d_ex_dy + 0.0f;
                }
                if (PZ && gzp) {
                    d_ey_dz = cpml4(d_ey_dz, splat(pbz), splat(paz), splat(pkz), q1z[0]);
                    d_ex_dz = cpml4(d_ex_dz, splat(pbz), splat(paz), splat(pkz), q1z[PZ ? 1 : 0]);
                } else {
                    d_ey_dz = d_ey_dz + 0.0f;
                    d_ex_dz = d_ex_dz + 0.0f;
                }
                if (PX && gxq) {
                    d_ez_dx = cpml4(d_ez_dx, pbx, pax, pkx, q1x[0]);
                    d_ey_dx = cpml4(d_ey_dx, pbx, pax, pkx, q1x[PX ? 1 : 0]);
                } else {
                    d_ez_dx = d_ez_dx + 0.0f;
                    d_ey_dx = d_ey_dx + 0.0f;
                }
            }
            h1x = madd4(da, cur.hx_o, -(db * (d_ez_dy - d_ey_dz)));
            h1y = madd4(da, cur.hy_o, -(db * (d_ex_dz - d_ez_dx)));
            h1z = madd4(da, cur.hz_o, -(db * (d_ey_dx - d_ex_dy)));
        }
        if (!live || !lp) {
            h1x = splat(0.f);
            h1y = splat(0.f);
            h1z = splat(0.f);
        }
        s_h1[0][me] = h1z;
        s_h1[1][me] = h1x;
        lds_barrier();
        // this row's E^n of the next plane, for the row below (read before
        // the next plane's first barrier, written after this one)
        s_ep[0][me] = cur.ez_n;
        s_ep[1][me] = cur.ex_n;

        // ---- step 1: E(n+1) at plane p (update_e_field[_pml]), sources,
        // probes ------------------------------------------------------------
        f4 e1x = splat(0.f), e1y = splat(0.f), e1z = splat(0.f);
        const float hz_l = __shfl_up(h1z.w, 1, 64), hy_l = __shfl_up(h1y.w, 1, 64);
        if (R >= 1 && it >= -1) {
            const f4 hz_my = s_h1[0][below], hx_my = s_h1[1][below];
            const f4 hz_mx = f4{hz_l, h1z.x, h1z.y, h1z.z};
            const f4 hy_mx = f4{hy_l, h1y.x, h1y.y, h1y.z};
            f4 ca, cb;
            coef4(s_ce, cur.ids, ca, cb);
            f4 d_hz_dy = h1z - hz_my, d_hy_dz = h1y - h1y_c;
            f4 d_hx_dz = h1x - h1x_c, d_hz_dx = h1z - hz_mx;
            f4 d_hy_dx = h1y - hy_mx, d_hx_dy = h1x - hx_my;
            if (PM && sh1) {
                if (PY && gyr) {
                    d_hz_dy = cpml4(d_hz_dy, splat(pby), splat(pay), splat(pky), q1y[PY ? 2 : 0]);
                    d_hx_dy = cpml4(d_hx_dy, splat(pby), splat(pay), splat(pky), q1y[PY ? 3 : 0]);
                } else {
                    d_hz_dy = d_hz_dy + 0.0f;
                    d_hx_dy = d_hx_dy + 0.0f;
                }
                if (PZ && gzp) {
                    d_hy_dz = cpml4(d_hy_dz, splat(pbz), splat(paz), splat(pkz), q1z[PZ ? 2 : 0]);
                    d_hx_dz = cpml4(d_hx_dz, splat(pbz), splat(paz), splat(pkz), q1z[PZ ? 3 : 0]);
                } else {
                    d_hy_dz = d_hy_dz + 0.0f;
                    d_hx_dz = d_hx_dz + 0.0f;
                }
                if (PX && gxq) {
                    d_hz_dx = cpml4(d_hz_dx, pbx, pax, pkx, q1x[PX ? 2 : 0]);
                    d_hy_dx = cpml4(d_hy_dx, pbx, pax, pkx, q1x[PX ? 3 : 0]);
                } else {
                    d_hz_dx = d_hz_dx + 0.0f;
                    d_hy_dx = d_hy_dx + 0.0f;
                }
            }
            e1x = madd4(ca, ex_c, cb * (d_hz_dy - d_hy_dz));
            e1y = madd4(ca, ey_c, cb * (d_hx_dz - d_hz_dx));
            e1z = madd4(ca, ez_c, cb * (d_hy_dx - d_hx_dy));
            if (live_row && lp) {
                const uint64_t key = (uint64_t)p * g.ny + y;
                if (point_row_flagged(pts.src, key)) {
                    for (int i = point_row_begin(pts.src, key); i < pts.src.n && pts.src.key[i] == key;
                         ++i) {
                        const PointEnt e = pts.src.ent[i];
                        const int k = (int)e.x - x;
                        if (live && k >= 0 && k < 4) {
                            const float v = sv[e.col];
                            if (e.comp == 0)
                                quad_source(e1x, k, v, e.hard);
                            else if (e.comp == 1)
                                quad_source(e1y, k, v, e.hard);
                            else
                                quad_source(e1z, k, v, e.hard);
                        }
                    }
                }
                if (own_row && it >= 0 && it < nit && point_row_flagged(pts.prb, key)) {
                    for (int i = point_row_begin(pts.prb, key); i < pts.prb.n && pts.prb.key[i] == key;
                         ++i) {
                        const PointEnt e = pts.prb.ent[i];
                        const int k = (int)e.x - x;
                        if (own && k >= 0 && k < 4) {
                            const f4 &f = e.comp == 0 ? e1x
                                        : e.comp == 1 ? e1y
                                        : e.comp == 2 ? e1z
                                        : e.comp == 3 ? h1x
                                        : e.comp == 4 ? h1y
                                                      : h1z;
                            pts.slot[e.col] = quad_get(f, k);
                        }
                    }
                }
            }
        }
        if (!live || !lp) {
            e1x = splat(0.f);
            e1y = splat(0.f);
            e1z = splat(0.f);
        }
        s_e1[it & 1][0][me] = e1z;
        s_e1[it & 1][1][me] = e1x;

        // ---- step 2: H(n+3/2) at plane p-1 ------------------------------
        // E(n+1) of plane p-1, as this row put it into LDS
        e1z_c = s_e1[(it - 1) & 1][0][me];
        e1x_c = s_e1[(it - 1) & 1][1][me];
        f4 h2x = splat(0.f), h2y = splat(0.f), h2z = splat(0.f);
        const float s2z = __shfl_down(e1z_c.x, 1, 64), s2y = __shfl_down(e1y_c.x, 1, 64);
        // the H-phase halves of the carried psi: previous plane's in, this one's out
        auto swap_in = [&](f4 *q, f4 *q1, int k0) {
            q[k0] = s_q[k0][me];
            q[k0 + 1] = s_q[k0 + 1][me];
            s_q[k0][me] = q1[k0];
            s_q[k0 + 1][me] = q1[k0 + 1];
        };
        if (QL) swap_in(qx, q1x, 0);
        if (QY) swap_in(qy, q1y, 0);
        if (QZ) swap_in(qz, q1z, 0);
        if (it >= 0 && R >= 1 && R <= RT - 2) {
            const int pb = (it - 1) & 1;
            const f4 ez_py = s_e1[pb][0][above], ex_py = s_e1[pb][1][above];
            const f4 ez_px2 = f4{e1z_c.y, e1z_c.z, e1z_c.w, s2z};
            const f4 ey_px2 = f4{e1y_c.y, e1y_c.z, e1y_c.w, s2y};
            f4 da, db;
            coef4(s_ch, ids_c, da, db);
            f4 d_ez_dy = ez_py - e1z_c, d_ey_dz = e1y - e1y_c;
            f4 d_ex_dz = e1x - e1x_c, d_ez_dx = ez_px2 - e1z_c;
            f4 d_ey_dx = ey_px2 - e1y_c, d_ex_dy = ex_py - e1x_c;
            if (PM && sh2) {
                if (PY && gyr) {
                    d_ez_dy = cpml4(d_ez_dy, splat(pby), splat(pay), splat(pky), qy[0]);
                    d_ex_dy = cpml4(d_ex_dy, splat(pby), splat(pay), splat(pky), qy[PY ? 1 : 0]);
                } else {
                    d_ez_dy = d_ez_dy + 0.0f;
                    d_ex_dy = d_ex_dy + 0.0f;
                }
                if (PZ && gz_c) {
                    d_ey_dz = cpml4(d_ey_dz, splat(pbz_c), splat(paz_c), splat(pkz_c), qz[0]);
                    d_ex_dz = cpml4(d_ex_dz, splat(pbz_c), splat(paz_c), splat(pkz_c), qz[PZ ? 1 : 0]);
                } else {
                    d_ey_dz = d_ey_dz + 0.0f;
                    d_ex_dz = d_ex_dz + 0.0f;
                }
                if (PX && gxq) {
                    d_ez_dx = cpml4(d_ez_dx, pbx, pax, pkx, qx[0]);
                    d_ey_dx = cpml4(d_ey_dx, pbx, pax, pkx, qx[PX ? 1 : 0]);
                } else {
                    d_ez_dx = d_ez_dx + 0.0f;
                    d_ey_dx = d_ey_dx + 0.0f;
                }
            }
            h2x = madd4(da, h1x_c, -(db * (d_ez_dy - d_ey_dz)));
            h2y = madd4(da, h1y_c, -(db * (d_ex_dz - d_ez_dx)));
            h2z = madd4(da, h1z_c, -(db * (d_ey_dx - d_ex_dy)));
            if (own && it >= 1) {  // the carried axis's H-phase psi(n+2)
                if (QL && gxq) {
                    const size_t o = (size_t)(p - 1) * psx + gxi;
                    st4(b.psi[0][0] + o, qx[0]);
                    st4(b.psi[0][1] + o, qx[1]);
                }
                if (QY && gyr) {
                    const size_t o = (size_t)(p - 1) * psy + gyo;
                    st4(b.psi[1][0] + o, qy[0]);
                    st4(b.psi[1][1] + o, qy[1]);
                }
                if (QZ && gz_c) {
                    const size_t o = (size_t)slab_index(g, 2, p - 1, g.nz) * plane + row;
                    st4(b.psi[2][0] + o, qz[0]);
                    st4(b.psi[2][1] + o, qz[1]);
                }
            }
        }
        if (!live || !lq) {
            h2x = splat(0.f);
            h2y = splat(0.f);
            h2z = splat(0.f);
        }
        s_h2[0][me] = h2z;
        s_h2[1][me] = h2x;
        lds_barrier();

        // ---- step 2: E(n+2) at plane p-1, stores of the owned cells -----
        const float h2z_l = __shfl_up(h2z.w, 1, 64), h2y_l = __shfl_up(h2y.w, 1, 64);
        if (QL) swap_in(qx, q1x, 2);
        if (QY) swap_in(qy, q1y, 2);
        if (QZ) swap_in(qz, q1z, 2);
        if (it >= 1 && R >= 2 && R <= RT - 2) {
            const f4 hz_my = s_h2[0][below], hx_my = s_h2[1][below];
            const f4 hz_mx = f4{h2z_l, h2z.x, h2z.y, h2z.z};
            const f4 hy_mx = f4{h2y_l, h2y.x, h2y.y, h2y.z};
            f4 ca, cb;
            coef4(s_ce, ids_c, ca, cb);
            f4 d_hz_dy = h2z - hz_my, d_hy_dz = h2y - h2y_m;
            f4 d_hx_dz = h2x - h2x_m, d_hz_dx = h2z - hz_mx;
            f4 d_hy_dx = h2y - hy_mx, d_hx_dy = h2x - hx_my;
            if (PM && sh2) {
                if (PY && gyr) {
                    d_hz_dy = cpml4(d_hz_dy, splat(pby), splat(pay), splat(pky), qy[PY ? 2 : 0]);
                    d_hx_dy = cpml4(d_hx_dy, splat(pby), splat(pay), splat(pky), qy[PY ? 3 : 0]);
                } else {
                    d_hz_dy = d_hz_dy + 0.0f;
                    d_hx_dy = d_hx_dy + 0.0f;
                }
                if (PZ && gz_c) {
                    d_hy_dz = cpml4(d_hy_dz, splat(pbz_c), splat(paz_c), splat(pkz_c), qz[PZ ? 2 : 0]);
                    d_hx_dz = cpml4(d_hx_dz, splat(pbz_c), splat(paz_c), splat(pkz_c), qz[PZ ? 3 : 0]);
                } else {
                    d_hy_dz = d_hy_dz + 0.0f;
                    d_hx_dz = d_hx_dz + 0.0f;
                }
                if (PX && gxq) {
                    d_hz_dx = cpml4(d_hz_dx, pbx, pax, pkx, qx[PX ? 2 : 0]);
                    d_hy_dx = cpml4(d_hy_dx, pbx, pax, pkx, qx[PX ? 3 : 0]);
                } else {
                    d_hz_dx = d_hz_dx + 0.0f;
                    d_hy_dx = d_hy_dx + 0.0f;
                }
            }
            const f4 e2x = madd4(ca, e1x_c, cb * (d_hz_dy - d_hy_dz));
            const f4 e2y = madd4(ca, e1y_c, cb * (d_hx_dz - d_hz_dx));
            const f4 e2z = madd4(ca, e1z_c, cb * (d_hy_dx - d_hx_dy));
            if (own) {
                const size_t o = (size_t)(p - 1) * plane + row;
                st4(b.f[0] + o, e2x);
                st4(b.f[1] + o, e2y);
                st4(b.f[2] + o, e2z);
                st4(b.f[3] + o, h2x);
                st4(b.f[4] + o, h2y);
                st4(b.f[5] + o, h2z);
                if (PX && gxq) {
                    for (int k = QL ? 2 : 0; k < NX; ++k)
                        st4(b.psi[0][k] + (size_t)(p - 1) * psx + gxi, qx[k]);
                }
                if (PY && gyr)
                    for (int k = QY ? 2 : 0; k < NY; ++k)
                        st4(b.psi[1][k] + (size_t)(p - 1) * psy + gyo, qy[k]);
                if (PZ && gz_c) {
                    const size_t oz = (size_t)slab_index(g, 2, p - 1, g.nz) * plane + row;
                    for (int k = QZ ? 2 : 0; k < NZ; ++k) st4(b.psi[2][k] + oz, qz[k]);
                }
            }
        }

        h2x_m = h2x;
        h2y_m = h2y;
        e1y_c = e1y;
        h1x_c = h1x;
        h1y_c = h1y;
        h1z_c = h1z;
        if (!QL)
            for (int k = 0; k < NX; ++k) qx[k] = q1x[k];
        if (!QY)
            for (int k = 0; k < NY; ++k) qy[k] = q1y[k];
        if (!QZ)
            for (int k = 0; k < NZ; ++k) qz[k] = q1z[k];
        shz_c = shz;
        gz_c = gzp;
        pbz_c = pbz;
        paz_c = paz;
        pkz_c = pkz;
        ids_c = cur.ids;
        ex_c = cur.ex_n;
        ey_c = cur.ey_n;
        ez_c = cur.ez_n;
    }
}

template <int PM, int LPRT>
__global__ __launch_bounds__(64 * twostep_waves<PM>()) void k_twostep4(Geo g, Ptrs a, Ptrs b,
                                                                       Box box,
                                                                       TwoStepPoints pts)
{
    twostep_tile<PM, LPRT, twostep_waves<PM>(), twostep_carry<PM>()>(
        g, a, b, box, pts, xcd_tile(blockIdx.x, box.gx * box.gy * box.gz));
}

#ifndef WF_STEP_FMA
// Batched two-step pass: the tiles of the regions of one variant over every
// scene of a batch in one launch (rows of box.lpr lanes); workgroup i
// computes tile t - box.first of item wg_item[t], t = i after XCD
// renumbering.
// The batch's x CPML variant <1> carries its x psi(n+1) in LDS (168 VGPRs)
// and runs 10 waves, three per SIMD: on the 96-scene set 1.823 against
// 1.878 ms per step with registers at 8 waves (LDS carry at 8 waves: 1.891).
#ifndef WF_BATCH_FYX
#define WF_BATCH_FYX 10
#endif
// The y and z variants <2>, <4> can carry theirs the same way
// (WF_BATCH_CARRY_YZ=1: 164-168 VGPRs at 10 waves); on the 96-scene set that
// measured 1.832 ms per step, the same as registers at 8 waves (182 VGPRs),
// which stay the default.  Registers at 10 waves spill.
#ifndef WF_BATCH_CARRY_YZ
#define WF_BATCH_CARRY_YZ 0
#endif
#ifndef WF_BATCH_FY1
#define WF_BATCH_FY1 (WF_BATCH_CARRY_YZ ? 10 : 8)
#endif
template <int PM>
constexpr int twostep_batch_waves()
{
    return PM == 0 ? kF2 : (PM == 1 ? WF_BATCH_FYX : ((PM == 2 || PM == 4) ? WF_BATCH_FY1 : kF2x));
}
// The axis whose psi(n+1) a batched variant carries in LDS (single-axis
// variants only: one axis's slots fit beside 10 waves' exchanges).
template <int PM>
constexpr int twostep_batch_carry()
{
    return PM == 1 ? 1 : ((WF_BATCH_CARRY_YZ && (PM == 2 || PM == 4)) ? PM : 0);
}

template <int PM>
__global__ __launch_bounds__(64 * twostep_batch_waves<PM>()) void k_twostep4_batch(
    const TwoStepItem *__restrict__ items, const uint32_t *__restrict__ wg_item, int total)
{
    const int t = xcd_tile(blockIdx.x, total);
    const TwoStepItem &it = items[wg_item[t]];
    twostep_tile<PM, 0, twostep_batch_waves<PM>(), twostep_batch_carry<PM>()>(
        it.g, it.a, it.b, it.box, it.pts, t - it.box.first);
}
#endif

}  // namespace

#ifndef WF_STEP_FMA
int fused_rows_per_tile(bool pml_or_ghost) { return (pml_or_ghost ? kFYx : kFY) - 1; }
int fused_rows_per_tile_x() { return WF_FUSED_XPML_FY - 1; }
int fused_rows_per_tile_ghost() { return fused_waves<0, true>() - 1; }
int split_rows_per_wg() { return kSY; }
int twostep_rows_per_tile(int pml_axes, int lpr)
{
    int waves = twostep_waves<7>();
    switch (pml_axes) {
    case 0: waves = twostep_waves<0>(); break;
    case 1: waves = twostep_waves<1>(); break;
    case 2: waves = twostep_waves<2>(); break;
    case 3: waves = twostep_waves<3>(); break;
    case 4: waves = twostep_waves<4>(); break;
    case 5: waves = twostep_waves<5>(); break;
    case 6: waves = twostep_waves<6>(); break;
    default: break;
    }
    return waves * (64 / lpr) - 3;
}
int twostep_batch_rows_per_tile(int pml_axes, int lpr)
{
    int waves = twostep_batch_waves<7>();
    switch (pml_axes) {
    case 0: waves = twostep_batch_waves<0>(); break;
    case 1: waves = twostep_batch_waves<1>(); break;
    case 2: waves = twostep_batch_waves<2>(); break;
    case 4: waves = twostep_batch_waves<4>(); break;
    default: break;
    }
    return waves * (64 / lpr) - 3;
}
#endif

hipError_t launch_twostep(const StepLaunch &s, const Box &box, const TwoStepPoints &pts,
                          hipStream_t st, int pml_axes)
{
    const int n = box.gx * box.gy * box.gz;
    if (n <= 0) return hipSuccess;
    const dim3 grid((unsigned)n);
#define WF_TWOSTEP_LAUNCH(PM, L)                                                                  \
    hipLaunchKernelGGL((k_twostep4<PM, L>), grid, dim3(64, twostep_waves<PM>()), 0, st, s.g, s.cur, \
                       s.alt, box, pts)
    if (pml_axes & 1) {  // x CPML regions: narrow rows, box.lpr lanes
        if (box.lpr < 3 || box.lpr > 32) return hipErrorInvalidValue;
        if (pml_axes == 1) WF_TWOSTEP_LAUNCH(1, 0);
        else if (pml_axes == 3) WF_TWOSTEP_LAUNCH(3, 0);
        else if (pml_axes == 5) WF_TWOSTEP_LAUNCH(5, 0);
        else if (pml_axes == 7) WF_TWOSTEP_LAUNCH(7, 0);
        else return hipErrorInvalidValue;
    } else if (pml_axes == 0) {
        if (box.lpr == 64) WF_TWOSTEP_LAUNCH(0, 64);
        else if (box.lpr == 32) WF_TWOSTEP_LAUNCH(0, 32);
        else return hipErrorInvalidValue;
    } else if (box.lpr == 32) {  // y and/or z CPML regions: rows of 32 or 64 lanes
        if (pml_axes == 2) WF_TWOSTEP_LAUNCH(2, 32);
        else if (pml_axes == 4) WF_TWOSTEP_LAUNCH(4, 32);
        else if (pml_axes == 6) WF_TWOSTEP_LAUNCH(6, 32);
        else return hipErrorInvalidValue;
    } else if (box.lpr == 64) {
        if (pml_axes == 2) WF_TWOSTEP_LAUNCH(2, 64);
        else if (pml_axes == 4) WF_TWOSTEP_LAUNCH(4, 64);
        else if (pml_axes == 6) WF_TWOSTEP_LAUNCH(6, 64);
        else return hipErrorInvalidValue;
    } else {
        return hipErrorInvalidValue;
    }
#undef WF_TWOSTEP_LAUNCH
    return hipGetLastError();
}

#ifndef WF_STEP_FMA
hipError_t launch_batch_twostep(const TwoStepItem *items, const uint32_t *wg_item, int total,
                                int pml_axes, hipStream_t st)
{
    if (total <= 0) return hipSuccess;
    const dim3 grid((unsigned)total);
#define WF_TWOSTEP_BATCH(PM)                                                                  \
    hipLaunchKernelGGL((k_twostep4_batch<PM>), grid, dim3(64, twostep_batch_waves<PM>()), 0, st, items, \
                       wg_item, total)
    switch (pml_axes) {
    case 0: WF_TWOSTEP_BATCH(0); break;
    case 1: WF_TWOSTEP_BATCH(1); break;
    case 2: WF_TWOSTEP_BATCH(2); break;
    case 3: WF_TWOSTEP_BATCH(3); break;
    case 4: WF_TWOSTEP_BATCH(4); break;
    case 5: WF_TWOSTEP_BATCH(5); break;
    case 6: WF_TWOSTEP_BATCH(6); break;
    case 7: WF_TWOSTEP_BATCH(7); break;
    default: return hipErrorInvalidValue;
    }
#undef WF_TWOSTEP_BATCH
    return hipGetLastError();
}
#endif

hipError_t launch_h(const StepLaunch &s, const BoxList &bl, hipStream_t st)
{
    if (bl.total == 0) return hipSuccess;
    const dim3 grid((unsigned)bl.total), block(64, kSY);
    if (s.ghost)
        hipLaunchKernelGGL((k_h4<true, true>), grid, block, 0, st, s.g, s.cur, s.alt, bl, s.zchunk);
    else if (s.pml)
        hipLaunchKernelGGL((k_h4<true, false>), grid, block, 0, st, s.g, s.cur, s.alt, bl, s.zchunk);
    else
        hipLaunchKernelGGL((k_h4<false, false>), grid, block, 0, st, s.g, s.cur, s.alt, bl, s.zchunk);
    return hipGetLastError();
}

hipError_t launch_e(const StepLaunch &s, const BoxList &bl, hipStream_t st)
{
    if (bl.total == 0) return hipSuccess;
    const dim3 grid((unsigned)bl.total), block(64, kSY);
    if (s.ghost)
        hipLaunchKernelGGL((k_e4<true, true>), grid, block, 0, st, s.g, s.alt, s.cur, s.alt, bl,
                           s.zchunk);
    else if (s.pml)
        hipLaunchKernelGGL((k_e4<true, false>), grid, block, 0, st, s.g, s.alt, s.cur, s.alt, bl,
                           s.zchunk);
    else
        hipLaunchKernelGGL((k_e4<false, false>), grid, block, 0, st, s.g, s.alt, s.cur, s.alt, bl,
                           s.zchunk);
    return hipGetLastError();
}

hipError_t launch_batch_h(const BatchScene *scenes, const uint32_t *wg_scene, int total,
                          int zchunk, int grid, hipStream_t st)
{
    if (total == 0) return hipSuccess;
    const int n = grid > 0 && grid < total ? grid : total;
    hipLaunchKernelGGL(k_h4_batch<true>, dim3((unsigned)n), dim3(64, kSY), 0, st, scenes,
                       wg_scene, total, zchunk);
    return hipGetLastError();
}

hipError_t launch_batch_e(const BatchScene *scenes, const uint32_t *wg_scene, int total,
                          int zchunk, int grid, hipStream_t st)
{
    if (total == 0) return hipSuccess;
    const int n = grid > 0 && grid < total ? grid : total;
    hipLaunchKernelGGL(k_e4_batch<true>, dim3((unsigned)n), dim3(64, kSY), 0, st, scenes,
                       wg_scene, total, zchunk);
    return hipGetLastError();
}

hipError_t launch_fused(const StepLaunch &s, const Box &box, bool pml_cells, hipStream_t st,
                        bool x_only)
{
    const int n = box.gx * box.gy * box.gz;
    if (n <= 0) return hipSuccess;
    if (pml_cells && s.ghost) return hipErrorInvalidValue;  // not a supported combination
    if (pml_cells && x_only)
        hipLaunchKernelGGL((k_fused4<2, false>), dim3((unsigned)n), dim3(64, WF_FUSED_XPML_FY), 0,
                           st, s.g, s.cur, s.alt, box);
    else if (pml_cells)
        hipLaunchKernelGGL((k_fused4<1, false>), dim3((unsigned)n), dim3(64, kFYx), 0, st, s.g,
                           s.cur, s.alt, box);
    else if (s.ghost)
        hipLaunchKernelGGL((k_fused4<0, true>), dim3((unsigned)n), dim3(64, fused_waves<0, true>()), 0, st, s.g,
                           s.cur, s.alt, box);
    else
        hipLaunchKernelGGL((k_fused4<0, false>), dim3((unsigned)n), dim3(64, kFY), 0, st, s.g,
                           s.cur, s.alt, box);
    return hipGetLastError();
}

#ifdef WF_STEP_FMA
}  // namespace fma
#endif
}  // namespace wf

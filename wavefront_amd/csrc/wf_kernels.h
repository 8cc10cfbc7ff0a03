// Device-side data structures and helpers shared by the HIP translation
// units (step kernels, point kernels, converters).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace wf {

// Grid geometry and boundary description, passed by value to kernels.
struct Geo {
    int nx, ny, nz;
    int nb[3];          // blocks per axis (n/8)
    int T[3];           // graded CPML depth per axis, voxels (0 = none)
    int sb[3];          // PML shell depth per axis, blocks (0 = none)
    unsigned periodic;  // bit a: axis a wraps
};

// Device pointers.  Fields are x-fastest SoA: index (z*ny + y)*nx + x.
// psi slabs hold only the non-identity CPML terms:
//   axis x: [(z*ny + y)*2Tx + gx],  gx in [0, 2Tx)
//   axis y: [(z*2Ty + gy)*nx + x]
//   axis z: [(gz*ny + y)*nx + x]
// where g = c for c < T and c - (n - 2T) for c >= n - T.
// Terms per axis a (H phase first, then E phase):
//   x: ez_dx, ey_dx, hz_dx, hy_dx
//   y: ez_dy, ex_dy, hz_dy, hx_dy
//   z: ey_dz, ex_dz, hy_dz, hx_dz
struct Ptrs {
    float *f[6];  // ex ey ez hx hy hz
    const uint8_t *id;
    const float4 *coef;       // 256 x (ca, cb, da, db), dequantized Q16.16
    const float *prof[3][3];  // [axis][b, a, inv_kappa], length n_axis
    float *psi[3][4];
    // Learned-boundary ghost planes (null = the zero boundary), per face
    // x_lo, x_hi, y_lo, y_hi, z_lo, z_hi and polarization: high faces hold
    // E just outside (x: Ey, Ez; y: Ez, Ex; z: Ex, Ey), low faces H (x: Hy, Hz;
    // y: Hz, Hx; z: Hx, Hy).  Planes are indexed x: z*ny + y, y: z*nx + x,
    // z: y*nx + x.
    const float *ghost[6][2];
};

// Source / probe tables.
struct DevSource {
    uint64_t off;  // voxel linear index
    int comp;      // 0..2
    int hard;
    uint32_t col;  // the source's column of values (values[row * ncol + col])
    uint32_t pad;
};

struct DevProbe {
    uint64_t off;
    int comp;  // 0..5
    int nfreq;
    int foff;  // offset into the flattened frequency axis
    int pad;
};

// Point work of the two-step kernel at its first step: sources (applied to
// E(n+1) wherever the kernel computes it) and probes (the values of owned
// cells recorded into one slot per probe).  Entries are sorted by row
// (z*ny + y), list order kept within a row; a bitmap flags the rows that
// hold any entry.
struct PointEnt {
    uint32_t x;
    uint16_t comp;  // sources: 0..2; probes: 0..5
    uint16_t hard;  // sources: '=' instead of '+='
    uint32_t col;   // sources: value column; probes: probe index
};

struct PointRows {
    const uint32_t *bits = nullptr;  // null: no entries
    const uint64_t *key = nullptr;   // row of each entry, ascending
    const PointEnt *ent = nullptr;
    int n = 0;
};

struct TwoStepPoints {
    PointRows src, prb;
    const float *values = nullptr;  // source values, values[row * ncol + col]
    int ncol = 0;
    uint64_t base = 0;               // step of the table's first row
    const uint64_t *step = nullptr;  // device step counter (the first step's loop index)
    float *slot = nullptr;           // probe values of the first step
};

// Four consecutive x-cells: the unit a lane loads and stores (16 B).
typedef float f4 __attribute__((ext_vector_type(4)));

// One-tap recursive convolution, src/fdtd.rs:136-163:
//   psi = b*psi + a*raw;  d = raw*inv_kappa + psi
__device__ __forceinline__ float cpml(float raw, float b, float a, float ik, float &psi)
{
    const float pn = b * psi + a * raw;
    psi = pn;
    return raw * ik + pn;
}

__device__ __forceinline__ f4 cpml4(f4 raw, f4 b, f4 a, f4 ik, f4 &psi)
{
    const f4 pn = b * psi + a * raw;
    psi = pn;
    return raw * ik + pn;
}

// A block (8 cells) on axis ax is in the PML shell (PmlAuxGrid::build,
// src/layout.rs:602-608).
__device__ __forceinline__ bool in_shell_axis(const Geo &g, int ax, int c)
{
    const int b = c >> 3;
    return g.sb[ax] > 0 && (b < g.sb[ax] || b >= g.nb[ax] - g.sb[ax]);
}

// Cell coordinate c lies in the graded layer of axis ax (non-identity profile).
__device__ __forceinline__ bool graded(const Geo &g, int ax, int c, int n)
{
    return g.T[ax] > 0 && (c < g.T[ax] || c >= n - g.T[ax]);
}

__device__ __forceinline__ int slab_index(const Geo &g, int ax, int c, int n)
{
    return c < g.T[ax] ? c : c - (n - 2 * g.T[ax]);
}

// Neighbour coordinate c+1 / c-1 on an axis of length n: -1 = outside the
// domain (reads the zero boundary, OUTER_ZERO_BLOCK in src/engine.rs:83).
__device__ __forceinline__ int plus1(int c, int n, bool wrap)
{
    return c + 1 < n ? c + 1 : (wrap ? 0 : -1);
}

__device__ __forceinline__ int minus1(int c, int n, bool wrap)
{
    return c > 0 ? c - 1 : (wrap ? n - 1 : -1);
}

// Workgroups are dispatched to the 8 XCDs round-robin; renumber them so that
// each XCD sweeps a contiguous run of tiles (x fastest, then y, then z-chunk)
// and the halo lines shared by neighbouring tiles stay in one XCD's L2.
__device__ __forceinline__ int xcd_tile(int i, int G)
{
    const int per = G >> 3;
    return i < (per << 3) ? (i & 7) * per + (i >> 3) : i;
}

// Workgroup barrier that orders LDS only: outstanding global loads (a
// prefetched plane) and stores are not waited for.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// A box of cells [x0,x1) x [y0,y1) x [z0,z1) and its tiling: lpr lanes per
// row (split kernels), gx*gy*gz tiles, the first of which has workgroup
// index `first` in its launch.  Fused kernel: tile column tx starts at
// x0 + tx * tstride (256: adjacent columns; larger: columns apart, cells
// between them not in the box).
struct Box {
    int x0, x1, y0, y1, z0, z1;
    int lpr;
    int gx, gy, gz;
    int first;
    int tstride = 256;
};

constexpr int kMaxBoxes = 6;
struct BoxList {
    Box b[kMaxBoxes];
    int n;
    int total;  // workgroups over all boxes
};

// Launch parameters for one timestep.
struct StepLaunch {
    Geo g;
    Ptrs cur;  // state at step n
    Ptrs alt;  // state n+1 (== cur for the in-place split variant)
    bool pml;
    int zchunk;
    hipStream_t stream;
    bool ghost = false;  // some face reads a learned ghost plane
};

// H kernel over the boxes: reads E^n, H^{n-1/2}, psi from cur; writes
// H(n+1/2) and the H-phase psi to alt.
hipError_t launch_h(const StepLaunch &s, const BoxList &bl, hipStream_t st);
// E kernel over the boxes: reads H(n+1/2) from alt and E^n / psi from cur;
// writes E(n+1) and the E-phase psi to alt.
hipError_t launch_e(const StepLaunch &s, const BoxList &bl, hipStream_t st);
// Single-pass kernel over a box, cur -> alt; tile (tx, ty, tz) marches the
// tz-th of box.gz even slices of the box's planes (s.zchunk is not used).
// pml_cells: the box may hold CPML shell cells (the kernel then updates
// their psi too); false selects the plain variant for a box wholly outside
// the shell (x_only: on the x axis only, the box's y and z ranges lying
// outside it).  s.ghost: E beyond the high faces comes from the learned ghost
// planes (with pml_cells false); the low faces' boundary E is left for the
// E kernel to redo.
hipError_t launch_fused(const StepLaunch &s, const Box &box, bool pml_cells, hipStream_t st,
                        bool x_only = false);
// Two-step kernel over a box: E(n+2), H(n+3/2) of the box's cells into
// s.alt from s.cur (E^n, H^{n-1/2}), the sources of step n+1 applied and
// its probes recorded on the way.  pml_axes (bit 1 x, 2 y, 4 z): the CPML
// axes of the cells the tiles compute, margins included (two cells below
// and one above the box on each axis, x: a quad); on the other axes those
// cells must lie outside the shell; cells outside the domain read as zero;
// no axis periodic.  Rows are box.lpr lanes wide (64 or 32 without an x
// CPML, 3..32 with one), box.lpr - 2 quads owned: tile column tx owns
// [x0 + tx tstride, + 4 (lpr - 2)) clipped to x1; box.gy row groups of
// twostep_rows_per_tile(pml_axes, lpr) rows; box.gz even z slices.
hipError_t launch_twostep(const StepLaunch &s, const Box &box, const TwoStepPoints &pts,
                          hipStream_t st, int pml_axes = 0);
int twostep_rows_per_tile(int pml_axes, int lpr);  // owned rows per tile
// Tile shapes of the two kernels (owned rows per fused tile -- of the plain
// variant, or of the CPML / ghost variants -- and rows per split workgroup
// of lpr = 64).
int fused_rows_per_tile(bool pml_or_ghost);
int fused_rows_per_tile_x();  // the x-only CPML variant
int fused_rows_per_tile_ghost();  // the ghost variant (learned faces)
int split_rows_per_wg();

// One scene of a batch: its geometry and state (updated in place by the
// split kernels over `box`, the whole scene), and the tables of its point
// work, all read by the batched kernels from device memory.
struct BatchScene {
    Geo g;
    Ptrs p;
    Box box;  // box.first = the scene's first workgroup in the batched launch
    const DevSource *src;     // grouped by (component, voxel), list order within a group
    const uint32_t *src_grp;  // group starts, nsrc_grp + 1 entries
    int nsrc_grp;
    const float *srcval;
    uint64_t src_base;
    const DevProbe *probe;
    const double *cos_tab, *sin_tab;
    const uint8_t *active;
    double *sums;
    unsigned long long *counts;
    float *trace;
    uint64_t probe_base;
    uint64_t *step;  // the scene's device step counter
    const float *slot;  // two-step passes: probe values of the first step
    int nsrc, nprobe, nfreq;
    int ncol;  // source value columns: values[row * ncol + src.col]
};

// One region of one scene in a batched two-step pass: the scene's geometry,
// the state read (a) and the state written (b), the region's box (rows of
// box.lpr lanes, 3..64; box.first = its first tile in its variant's launch)
// and the point work of the pass's first step.
struct TwoStepItem {
    Geo g;
    Ptrs a, b;
    Box box;
    TwoStepPoints pts;
};

// Two-step kernel of the CPML variant pml_axes over the items' tiles in one
// launch: tile t (after XCD renumbering) belongs to item wg_item[t].  Every
// item's box obeys launch_twostep's rules for that variant, with any row
// width.
hipError_t launch_batch_twostep(const TwoStepItem *items, const uint32_t *wg_item, int total,
                                int pml_axes, hipStream_t st);
int twostep_batch_rows_per_tile(int pml_axes, int lpr);  // owned rows per batched tile

// H / E kernels over every scene of a batch in one launch: tile t (after
// XCD renumbering) belongs to scene wg_scene[t].  grid = 0: one workgroup
// per tile; otherwise `grid` persistent workgroups stride over the tiles.
hipError_t launch_batch_h(const BatchScene *scenes, const uint32_t *wg_scene, int total,
                          int zchunk, int grid, hipStream_t st);
hipError_t launch_batch_e(const BatchScene *scenes, const uint32_t *wg_scene, int total,
                          int zchunk, int grid, hipStream_t st);

// The same step kernels built with FP contraction (fused multiply-adds).
namespace fma {
hipError_t launch_h(const StepLaunch &s, const BoxList &bl, hipStream_t st);
hipError_t launch_e(const StepLaunch &s, const BoxList &bl, hipStream_t st);
hipError_t launch_fused(const StepLaunch &s, const Box &box, bool pml_cells, hipStream_t st,
                        bool x_only = false);
hipError_t launch_twostep(const StepLaunch &s, const Box &box, const TwoStepPoints &pts,
                          hipStream_t st, int pml_axes = 0);
hipError_t launch_batch_h(const BatchScene *scenes, const uint32_t *wg_scene, int total,
                          int zchunk, int grid, hipStream_t st);
hipError_t launch_batch_e(const BatchScene *scenes, const uint32_t *wg_scene, int total,
                          int zchunk, int grid, hipStream_t st);
}  // namespace fma

}  // namespace wf

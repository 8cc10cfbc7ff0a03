// Device engine for gfx950: context management, the H/E timestep kernels,
// source/probe point kernels and the SoA <-> reference-AoSoA converters.
//
// Built with -ffp-contract=off: every f32 operation in the update kernels is
// a separately rounded IEEE operation in the same order as the reference's
// src/fdtd.rs, so results are bit-identical to the CPU engine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "wf_internal.h"
#include "wf_kernels.h"
#include "wf_learned.h"
#include "wf_writer.h"

using namespace wf;

#define WF_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(WF_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                           \
    } while (0)

namespace {

// ---------------------------------------------------------------------------
// Plain device copy of known size: the practical HBM ceiling and the
// reference point that scales the FETCH_SIZE / WRITE_SIZE counters.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_copy(const float4 *__restrict__ src,
                                              float4 *__restrict__ dst, uint64_t n4)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n4) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Sources (src/source.rs:79-95).  The reference applies them in list order;
// only sources on the same field value can see each other, so the table is
// kept grouped by (component, voxel), each group in list order, and one lane
// applies one group: the same result as the sequential loop.
// ---------------------------------------------------------------------------
// `row` is the device step counter minus the step at which the table was
// set, so a captured graph of these launches stays valid for any step.
__device__ __forceinline__ void apply_source(const Ptrs &p, const DevSource &q,
                                             const float *__restrict__ v)
{
    float *f = p.f[q.comp] + q.off;
    if (q.hard)
        *f = v[q.col];
    else
        *f = *f + v[q.col];
}

// One workgroup's lanes over the groups [grp[g], grp[g+1]) of a scene's
// grouped table, each lane a group in list order.
__device__ __forceinline__ void apply_source_groups(const Ptrs &p, const DevSource *__restrict__ src,
                                                   const uint32_t *__restrict__ grp, int ngrp,
                                                   int ncol, const float *__restrict__ values,
                                                   uint64_t row)
{
    const float *v = values + row * (uint64_t)ncol;
    for (int g = threadIdx.x; g < ngrp; g += blockDim.x)
        for (uint32_t s = grp[g]; s < grp[g + 1]; ++s) apply_source(p, src[s], v);
}

// One lane per group [grp[i], grp[i+1]) of the grouped table; a step's
// values are one row of ncol columns.
__global__ void k_sources(Ptrs p, const DevSource *__restrict__ src, int ncol,
                          const uint32_t *__restrict__ grp, int ngrp,
                          const float *__restrict__ values, const uint64_t *__restrict__ step,
                          uint64_t base)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ngrp) return;
    const float *v = values + (*step - base) * (uint64_t)ncol;
    for (uint32_t s = grp[i]; s < grp[i + 1]; ++s) apply_source(p, src[s], v);
}

// ---------------------------------------------------------------------------
// Probes (src/probe.rs:126-138): one lane per probe, frequencies in order.
// ---------------------------------------------------------------------------
// slot (two-step passes, first step): the values the two-step kernels
// recorded, one per probe.
__device__ __forceinline__ void accumulate_probe(const Ptrs &p, const DevProbe &q, int i, int n,
                                                 int nf, const double *__restrict__ cos_tab,
                                                 const double *__restrict__ sin_tab,
                                                 const uint8_t *__restrict__ active,
                                                 uint64_t row, double *sums,
                                                 unsigned long long *counts, float *trace,
                                                 const float *slot = nullptr)
{
    const float v32 = slot ? slot[i] : p.f[q.comp][q.off];
    if (trace) trace[row * (uint64_t)n + i] = v32;
    if (!active[row * (uint64_t)n + i]) return;
    const double v = (double)v32;
    const double *ct = cos_tab + row * (uint64_t)nf + q.foff;
    const double *st = sin_tab + row * (uint64_t)nf + q.foff;
    double *sc = sums + q.foff;
    double *ss = sums + nf + q.foff;
    for (int j = 0; j < q.nfreq; ++j) {
        sc[j] = sc[j] + v * ct[j];
        ss[j] = ss[j] + v * st[j];
    }
    counts[i] += 1ull;
}

__global__ void k_probes(Ptrs p, const DevProbe *__restrict__ pr, int n, int nf,
                         const double *__restrict__ cos_tab, const double *__restrict__ sin_tab,
                         const uint8_t *__restrict__ active, const uint64_t *__restrict__ step,
                         uint64_t base, double *sums, unsigned long long *counts, float *trace,
                         const float *slot)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    accumulate_probe(p, pr[i], i, n, nf, cos_tab, sin_tab, active, *step - base, sums, counts,
                     trace, slot);
}

// Batched point work, one workgroup per scene: sources, then (next launch)
// probes and the scene's step counter.
__global__ void k_sources_batch(const BatchScene *__restrict__ sc)
{
    const BatchScene &d = sc[blockIdx.x];
    if (d.nsrc == 0) return;
    apply_source_groups(d.p, d.src, d.src_grp, d.nsrc_grp, d.ncol, d.srcval, *d.step - d.src_base);
}

__global__ void k_probes_batch(const BatchScene *__restrict__ sc)
{
    const BatchScene &d = sc[blockIdx.x];
    const uint64_t step = *d.step;
    for (int i = threadIdx.x; i < d.nprobe; i += blockDim.x)
        accumulate_probe(d.p, d.probe[i], i, d.nprobe, d.nfreq, d.cos_tab, d.sin_tab, d.active,
                         step - d.probe_base, d.sums, d.counts, d.trace);
    __syncthreads();  // every lane has read the counter
    if (threadIdx.x == 0) *d.step = step + 1ull;
}

// Point work of a batched two-step pass, one workgroup per scene (d.p: the
// state the pass wrote): the first step's probes from the values the
// two-step kernels recorded, the second step's sources (list order) and
// probes, the step counter advanced by two.
__global__ void k_pair_points_batch(const BatchScene *__restrict__ sc)
{
    const BatchScene &d = sc[blockIdx.x];
    const uint64_t step = *d.step;
    for (int i = threadIdx.x; i < d.nprobe; i += blockDim.x)
        accumulate_probe(d.p, d.probe[i], i, d.nprobe, d.nfreq, d.cos_tab, d.sin_tab, d.active,
                         step - d.probe_base, d.sums, d.counts, d.trace, d.slot);
    if (d.nsrc)
        apply_source_groups(d.p, d.src, d.src_grp, d.nsrc_grp, d.ncol, d.srcval,
                            step + 1 - d.src_base);
    __syncthreads();  // the second step's sources are in place
    for (int i = threadIdx.x; i < d.nprobe; i += blockDim.x)
        accumulate_probe(d.p, d.probe[i], i, d.nprobe, d.nfreq, d.cos_tab, d.sin_tab, d.active,
                         step + 1 - d.probe_base, d.sums, d.counts, d.trace, nullptr);
    __syncthreads();  // every lane has read the counter
    if (threadIdx.x == 0) *d.step = step + 2ull;
}

// Advances the device step counter at the end of every step.
__global__ void k_tick(uint64_t *step)
{
    if (threadIdx.x == 0) *step += 1ull;
}

// ---------------------------------------------------------------------------
// SoA <-> reference AoSoA (src/engine.rs:622-640) for block planes
// [bz0, bz0 + nbz_chunk).  One thread per voxel of the chunk.
// ---------------------------------------------------------------------------
__global__ void k_pack(Ptrs p, Geo g, int bz0, uint64_t nvox, float *__restrict__ out)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nvox) return;
    const uint64_t blk = t >> 9;
    const int l = (int)(t & 511);
    const int lx = l & 7, ly = (l >> 3) & 7, lz = l >> 6;
    const int bx = (int)(blk % g.nb[0]);
    const int by = (int)((blk / g.nb[0]) % g.nb[1]);
    const int bz = bz0 + (int)(blk / ((uint64_t)g.nb[0] * g.nb[1]));
    const size_t o = ((size_t)(bz * 8 + lz) * g.ny + (by * 8 + ly)) * (size_t)g.nx + bx * 8 + lx;
    float *dst = out + blk * 3072 + l;
#pragma unroll
    for (int c = 0; c < 6; ++c) dst[c * 512] = p.f[c][o];
}

__global__ void k_unpack(Ptrs p, Geo g, int bz0, uint64_t nvox, const float *__restrict__ in)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nvox) return;
    const uint64_t blk = t >> 9;
    const int l = (int)(t & 511);
    const int lx = l & 7, ly = (l >> 3) & 7, lz = l >> 6;
    const int bx = (int)(blk % g.nb[0]);
    const int by = (int)((blk / g.nb[0]) % g.nb[1]);
    const int bz = bz0 + (int)(blk / ((uint64_t)g.nb[0] * g.nb[1]));
    const size_t o = ((size_t)(bz * 8 + lz) * g.ny + (by * 8 + ly)) * (size_t)g.nx + bx * 8 + lx;
    const float *src = in + blk * 3072 + l;
#pragma unroll
    for (int c = 0; c < 6; ++c) p.f[c][o] = src[c * 512];
}

// psi in the reference PmlAux layout (src/layout.rs:545-558) for the shell
// blocks listed in `blocks` (packed bx | by<<20 | bz<<40).  Terms whose axis
// is not graded at a voxel are identically +0 in the reference.
__global__ void k_pack_psi(Ptrs p, Geo g, const uint64_t *__restrict__ blocks, uint64_t nblk,
                           float *__restrict__ out)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nblk * 512) return;
    const uint64_t b = t >> 9;
    const int l = (int)(t & 511);
    const uint64_t code = blocks[b];
    const int bx = (int)(code & 0xFFFFF), by = (int)((code >> 20) & 0xFFFFF),
              bz = (int)(code >> 40);
    const int x = bx * 8 + (l & 7), y = by * 8 + ((l >> 3) & 7), z = bz * 8 + (l >> 6);
    const bool gx = graded(g, 0, x, g.nx), gy = graded(g, 1, y, g.ny), gz = graded(g, 2, z, g.nz);
    const size_t nx = g.nx;
    const size_t ix = gx ? ((size_t)z * g.ny + y) * (2 * g.T[0]) + slab_index(g, 0, x, g.nx) : 0;
    const size_t iy = gy ? ((size_t)z * (2 * g.T[1]) + slab_index(g, 1, y, g.ny)) * nx + x : 0;
    const size_t iz = gz ? ((size_t)slab_index(g, 2, z, g.nz) * g.ny + y) * nx + x : 0;
    // PmlAux order -> (axis, term)
    const int ax[12] = {0, 0, 1, 1, 2, 2, 0, 0, 1, 1, 2, 2};
    const int tm[12] = {0, 1, 0, 1, 0, 1, 2, 3, 3, 2, 2, 3};
    float *dst = out + b * 6144 + l;
    for (int k = 0; k < 12; ++k) {
        const int a = ax[k];
        const bool on = a == 0 ? gx : (a == 1 ? gy : gz);
        const size_t i = a == 0 ? ix : (a == 1 ? iy : iz);
        dst[k * 512] = on ? p.psi[a][tm[k]][i] : 0.0f;
    }
}

}  // namespace

// ===========================================================================
// Context
// ===========================================================================
// One box of a two-step pass run through the batched kernel: its scene,
// the box (box.first: its first tile in its variant's launch) and the
// variant's CPML axes.
struct PassItem {
    int scene;
    wf::Box box;
    int axes;
};

struct wf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    Geo g{};
    Ptrs p{};     // current state (fields + psi); sources, probes, I/O use it
    Ptrs alt{};   // second state for the fused kernel (written, then swapped)
    bool pml = false;
    bool fused = true;
    int zchunk = 32;    // planes per workgroup: split variant, low-face fix-up boxes
    int fzchunk = 1;    // planes per workgroup of the frame kernels (fused variant)
    bool zchunk_env = false;  // WF_ZCHUNK given: the fused kernel takes it too
    int cus = 256;      // compute units of the device
    Box interior{};     // fused variant: cells swept by the single-pass kernel
    BoxList frame{};    // fused variant: the rest (holds the whole CPML shell)
    // fused variant, x axis folded into the single pass (plan_boxes): the
    // interior's middle columns (plain kernel), its two edge columns with the
    // x CPML cells (x-only CPML kernel), and the frame without x slabs; used
    // while no face hook is set
    bool xfold = false;
    bool xfold_env = true;  // WF_XFOLD=0: never fold
    Box xmid{}, xedges{};  // xedges: two tile columns, x from 0 and from nx - 256
    BoxList framex{};
    Box all{};          // fused variant with a CPML: the whole domain, one pass
    Box allg{};         // the same swept by the ghost variant (learned faces)
    BoxList lofix{};    // learned faces without a CPML: the low faces' boundary cells
    bool pml_fused = false;  // WF_PML_FUSED=1: single pass over the shell too (slower, see launch_step)
    BoxList whole{};    // split variant: the whole domain
    // two steps per pass (launch_pair): the domain cut into regions, shell
    // below / interior / shell above on each axis, each swept by the
    // two-step variant of its CPML axes (bit 1 x, 2 y, 4 z): per y and z range
    // the x interior range and the two x shell ranges (one launch, tile
    // columns nx - W apart); the point tables of the kernels' first step
    bool pair_env = true;   // WF_TWOSTEP=0: single steps only
    bool pair_geo = false;  // plan_boxes: the geometry allows it
    struct Region {
        Box box;
        int axes;
    };
    Region regions[18]{};
    int nregions = 0;
    struct PointTable {
        uint32_t *bits = nullptr;
        uint64_t *key = nullptr;
        PointEnt *ent = nullptr;
        int n = 0;
    } src_pts, prb_pts;
    float *d_pslot = nullptr;   // probe values of a pass's first step, per probe
    // schedule: 0 the regions on the two-step kernels (18 launches per pass
    // at most), 1 rows through the batched two-step kernel (full x rows per
    // y and z range, one launch per CPML variant), 2 single steps; chosen at
    // wf_create by timing for grids under 2^25 cells, else 0 (WF_SCHEDULE
    // sets it)
    int sched = 0;
    double sched_ms[3] = {-1.0, -1.0, -1.0};
    std::vector<PassItem> rows_plan;
    int rows_first[8] = {}, rows_total[8] = {};
    uint64_t rows_bytes = 0;
    TwoStepItem *d_ritems = nullptr;  // [2][rows_plan.size()]: the state in c->p at
    uint32_t *d_rwg = nullptr;        // writing (rows_key) read, then the other
    const float *rows_key = nullptr;
    bool rows_dirty = true;
    hipStream_t side = nullptr;            // frame H kernel, overlapped
    hipEvent_t fork = nullptr, join = nullptr;
    bool overlap = false;                  // WF_OVERLAP=1
    bool hook_overlap = true;              // WF_HOOK_OVERLAP=0 serializes the face hooks

    float *d_fields = nullptr;
    float *d_fields2 = nullptr;
    uint64_t fstride = 0;  // floats from one field of a state to the next (>= cells)
    float *d_psi2 = nullptr;
    uint8_t *d_id = nullptr;
    float4 *d_coef = nullptr;
    float *d_prof = nullptr;
    float *d_psi = nullptr;
    size_t psi_bytes = 0;       // each of d_psi, d_psi2
    uint64_t nshell_blocks = 0;

    // field placements timed at wf_create (ms per step) and the one kept
    double place_ms[17] = {};
    int place_n = 0, place_kept = 0;
    uint64_t *d_shell_list = nullptr;

    // sources
    uint32_t nsrc = 0;
    uint32_t nsrc_col = 0;  // source value columns (a row per step)
    DevSource *d_src = nullptr;     // grouped by (component, voxel), list order within a group
    uint32_t *d_src_grp = nullptr;  // group starts, ngrp + 1 entries
    uint32_t nsrc_grp = 0;
    float *d_srcval = nullptr;
    uint64_t src_base = 0, src_steps = 0;

    // probes
    uint32_t nprobe = 0, nfreq = 0;
    DevProbe *d_probe = nullptr;
    double *d_cos = nullptr, *d_sin = nullptr, *d_sums = nullptr;
    uint8_t *d_active = nullptr;
    unsigned long long *d_counts = nullptr;
    float *d_trace = nullptr;
    uint64_t probe_base = 0, probe_steps = 0;

    uint64_t step = 0;
    uint64_t *d_step = nullptr;          // device copy of `step`, read by point kernels
    hipGraphExec_t gexec = nullptr;      // kGraphSteps captured steps
    const float *gstate = nullptr;       // p.f[0] when the graph was captured
    bool use_graph = true;               // WF_GRAPH=0 disables

    float *d_stage = nullptr;
    uint64_t stage_bytes = 0;

    // wf_snapshot_async: a device copy of the packed state and its copy-out
    float *d_snap = nullptr;
    hipStream_t snap_stream = nullptr;
    hipEvent_t snap_packed = nullptr, snap_done = nullptr;
    bool snap_pending = false;

    // learned boundary: face hooks (predict or record) and their buffers,
    // per face x_lo, x_hi, y_lo, y_hi, z_lo, z_hi
    bool hook_on[6] = {false, false, false, false, false, false};
    FaceHook hook[6] = {};
    int face_passive = 1;      // predicted faces: passive limit on (wf_set_face_limits)
    float *d_pool[6] = {};     // limited faces: energy credit and flux partials
    float face_retain = 0.5f;  // share of the absorbed energy a limited face keeps
    float face_gamma = 8.0f;   // and the magnitude bound's factor (0 = none)
    float *d_ghost[6] = {};     // [2][nslow][nfast]
    float *d_mem[6] = {};       // [2][npos][4]
    float *d_samples[6] = {};   // [rows][2][npos][kSampleStride]
    MlpPacked *d_net[6] = {};
    bool ghost_reads = false;   // some predicted face feeds the step kernels

    // step kernel set: exact (contraction off, bit-identical to the CPU
    // engine) or contracted (fused multiply-adds)
    bool contract = false;
    hipError_t (*lh)(const StepLaunch &, const BoxList &, hipStream_t) = launch_h;
    hipError_t (*le)(const StepLaunch &, const BoxList &, hipStream_t) = launch_e;
    hipError_t (*lf)(const StepLaunch &, const Box &, bool, hipStream_t, bool) = launch_fused;
    hipError_t (*lt)(const StepLaunch &, const Box &, const TwoStepPoints &, hipStream_t, int) =
        launch_twostep;
};

namespace {

template <typename T>
int dmalloc(T **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) return WF_OK;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(p), bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(e == hipErrorOutOfMemory ? WF_ERR_OOM : WF_ERR_HIP,
                    "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    }
    return WF_OK;
}

template <typename T>
void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

// A captured step graph holds the table pointers: drop it with the tables.
void drop_graph_exec(wf_ctx *c)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
}

void free_points(wf_ctx::PointTable &t)
{
    dfree(t.bits);
    dfree(t.key);
    dfree(t.ent);
    t.n = 0;
}

// Row table of points for the two-step kernel: entries sorted by row
// (z*ny + y) keeping their given order within a row, and the bitmap of
// rows holding any.
int upload_points(const wf_ctx *c, std::vector<std::pair<uint64_t, PointEnt>> &v,
                  wf_ctx::PointTable &t)
{
    if (v.empty()) return WF_OK;
    std::stable_sort(v.begin(), v.end(),
                     [](const std::pair<uint64_t, PointEnt> &a, const std::pair<uint64_t, PointEnt> &b) {
                         return a.first < b.first;
                     });
    const uint64_t rows = (uint64_t)c->g.ny * c->g.nz;
    std::vector<uint32_t> bits((rows + 31) / 32, 0u);
    std::vector<uint64_t> key(v.size());
    std::vector<PointEnt> ent(v.size());
    for (size_t i = 0; i < v.size(); ++i) {
        key[i] = v[i].first;
        ent[i] = v[i].second;
        bits[key[i] >> 5] |= 1u << (key[i] & 31);
    }
    if (int rc = dmalloc(&t.bits, bits.size() * sizeof(uint32_t))) return rc;
    if (int rc = dmalloc(&t.key, key.size() * sizeof(uint64_t))) return rc;
    if (int rc = dmalloc(&t.ent, ent.size() * sizeof(PointEnt))) return rc;
    WF_HIP(hipMemcpy(t.bits, bits.data(), bits.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    WF_HIP(hipMemcpy(t.key, key.data(), key.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    WF_HIP(hipMemcpy(t.ent, ent.data(), ent.size() * sizeof(PointEnt), hipMemcpyHostToDevice));
    t.n = (int)v.size();
    return WF_OK;
}

void free_sources(wf_ctx *c)
{
    drop_graph_exec(c);
    dfree(c->d_src);
    dfree(c->d_src_grp);
    c->nsrc_grp = 0;
    dfree(c->d_srcval);
    free_points(c->src_pts);
    c->rows_dirty = true;  // the batched kernel's item tables hold the point tables
    c->nsrc = 0;
    c->nsrc_col = 0;
    c->src_steps = 0;
}

void free_probes(wf_ctx *c)
{
    drop_graph_exec(c);
    dfree(c->d_probe);
    dfree(c->d_cos);
    dfree(c->d_sin);
    dfree(c->d_sums);
    dfree(c->d_active);
    dfree(c->d_counts);
    dfree(c->d_trace);
    free_points(c->prb_pts);
    dfree(c->d_pslot);
    c->rows_dirty = true;
    c->nprobe = 0;
    c->nfreq = 0;
    c->probe_steps = 0;
}

// Tiles a box for the split kernels: lanes per row = the smallest power of
// two covering the box's quads (at most lpr_cap <= 64), kSY waves of 64/lpr
// rows each.  lpr_cap = 0 fits the rows instead: a row of q quads takes
// ceil(q / ceil(q / 64)) lanes (q itself up to 64), floor(64 / lpr) rows per
// wave, so narrow rows leave few lanes idle (40 cells: 10 lanes, 6 rows
// and 4 idle lanes per wave instead of 16-lane rows with 6 idle each).
Box split_box(int x0, int x1, int y0, int y1, int z0, int z1, int zchunk, int lpr_cap = 64)
{
    Box b{x0, x1, y0, y1, z0, z1, 64, 0, 0, 0, 0};
    const int quads = (x1 - x0 + 3) / 4;
    int lpr = 1;
    if (lpr_cap <= 0) {
        const int segs = (std::max(quads, 1) + 63) / 64;
        lpr = (std::max(quads, 1) + segs - 1) / segs;
    } else {
        while (lpr < quads && lpr < lpr_cap) lpr *= 2;
    }
    b.lpr = lpr;
    const int rows = split_rows_per_wg() * (64 / lpr);
    b.gx = (quads + lpr - 1) / lpr;
    b.gy = (y1 - y0 + rows - 1) / rows;
    b.gz = (z1 - z0 + zchunk - 1) / zchunk;
    return b;
}

void add_box(BoxList &bl, const Box &b)
{
    if (b.x1 <= b.x0 || b.y1 <= b.y0 || b.z1 <= b.z0) return;
    bl.b[bl.n] = b;
    bl.b[bl.n].first = bl.total;
    bl.total += b.gx * b.gy * b.gz;
    bl.n += 1;
}

// Interior box of the fused kernel: on every axis with a CPML layer of T
// cells, the cells from T+1 (T+4 on x, to keep quads aligned) to n-T, so
// that neither its cells nor its one-cell halos (x-1, y-1, z-1) are in the
// shell; the rest of the domain becomes up to six frame boxes.
// Number of z slices of a box for the fused kernel, which runs one
// workgroup per CU at a time (12 waves at 164 VGPRs, or 8 at up to 232):
// the launch takes about ceil(tiles * slices / CUs) rounds of (planes per
// slice + the `extra` planes swept before and after it: 1 for the fused
// kernel, 3 for the two-step kernel), so pick the slice count that
// minimises that product (ties: more slices).  At 1024^3 (356 tiles, 975 planes, 256 CUs)
// that is 5 slices of 195 planes, 8.09 ms per launch against 8.17 ms for
// 31 slices of <= 32 planes (whose last round is a tenth full); at 312^3
// with a 32-cell CPML 11 slices, 0.179 against 0.187 ms.
int fused_slices(long tiles, int planes, int cus, int extra = 1)
{
    if (planes <= 0 || tiles <= 0) return 1;
    long best_cost = -1;
    int best = 1;
    for (int gz = 1; gz <= planes; ++gz) {
        const long rounds = (tiles * gz + cus - 1) / cus;
        const long cost = rounds * ((planes + gz - 1) / gz + extra);
        if (best_cost < 0 || cost <= best_cost) {
            best_cost = cost;
            best = gz;
        }
    }
    return best;
}

// Algorithmic bytes of one two-step pass over a box's owned cells (tile
// column tx owns [x0 + tx tstride, + 4 (lpr - 2)) clipped to x1): 49 B per
// cell for the two steps, and 32 B per graded axis of a cell (its four
// non-identity psi terms, each read once and written once).
uint64_t twostep_box_bytes(const Geo &g, const Box &b)
{
    auto graded_at = [&](int a, int c, int n) {
        return (uint64_t)(g.T[a] > 0 && (c < g.T[a] || c >= n - g.T[a]));
    };
    uint64_t wx = 0, gx = 0, gy = 0, gz = 0;
    for (int tx = 0; tx < b.gx; ++tx) {
        const int x0 = b.x0 + tx * b.tstride, x1 = std::min(x0 + 4 * (b.lpr - 2), b.x1);
        for (int x = x0; x < x1; ++x) {
            ++wx;
            gx += graded_at(0, x, g.nx);
        }
    }
    for (int y = b.y0; y < b.y1; ++y) gy += graded_at(1, y, g.ny);
    for (int z = b.z0; z < b.z1; ++z) gz += graded_at(2, z, g.nz);
    const uint64_t wy = b.y1 - b.y0, wz = b.z1 - b.z0;
    return 49ull * wx * wy * wz + 32ull * (gx * wy * wz + gy * wx * wz + gz * wx * wy);
}

// Tile columns of a box whose x range holds q quads, for the two-step
// variant `axes`, rows of at most `cap` lanes: the column count (from the
// fewest possible up to three more) whose columns x row groups are fewest,
// each column as narrow as the count allows (ties: fewer columns).  Short
// rows leave lanes idle (64 / lanes-per-row rows per wave).
void fit_columns(Box &b, int cap, int axes)
{
    const int q = std::max((b.x1 - b.x0) / 4, 1), h = b.y1 - b.y0;
    const int k0 = (q + cap - 3) / (cap - 2);
    long best = -1;
    for (int k = k0; k <= std::min(q, k0 + 3); ++k) {
        const int per = (q + k - 1) / k;
        const int rows = twostep_batch_rows_per_tile(axes, per + 2);
        const long tiles = (long)k * ((h + rows - 1) / rows);
        if (best < 0 || tiles < best) {
            best = tiles;
            b.lpr = per + 2;
            b.tstride = 4 * per;
            b.gx = k;
        }
    }
}

// Two-step items of one scene (mode -1: full rows unless the x interior
// range is at least 496 cells wide, then the region split; 0: the region
// split; 1: full rows; 2: the whole scene in one box):
//  * the region split, as the solo plan (plan_boxes): per y and z range
//    (shell below / interior / shell above) the x interior range, rows
//    fitted to its width (64 lanes at most in the interior, 32 in the y and
//    z shell ranges), and the two x shell ranges as narrow columns;
//  * full rows: per y and z range one box over every x, on the variant
//    with the x CPML (rows of up to 32 lanes).  The narrow x shell columns
//    waste lanes and lines; on the 96-scene set (edges 32-192) full rows
//    took 1.89 ms per step against 2.54 for the split and 2.31 for whole-
//    scene boxes, measured on MI355X;
//  * the whole scene as one box of the every-axis CPML variant: where the
//    split does not fit (x interior under 8 cells, or y or z shells that
//    meet), and as mode 2.
void pass_items(const Geo &g, int mode, int scene, std::vector<PassItem> &items)
{
    const int n[3] = {g.nx, g.ny, g.nz};
    const int Wx = g.T[0] + 4, lprx = Wx / 4 + 2;
    int tl[3], th[3];
    tl[0] = Wx;
    th[0] = n[0] - Wx;
    for (int a = 1; a < 3; ++a) {
        tl[a] = g.T[a] + 2;
        th[a] = n[a] - std::max(g.T[a], 1) - 1;
    }
    const bool fits = lprx <= 32 && th[0] - tl[0] >= 8 && th[1] >= tl[1] && th[2] >= tl[2];
    if (mode < 0) mode = th[0] - tl[0] >= 496 ? 0 : 1;
    if (!fits || mode == 2) {
        Box w{0, n[0], 0, n[1], 0, n[2], 32, 0, 0, 0, 0};
        fit_columns(w, 32, 7);
        items.push_back({scene, w, 7});
        return;
    }
    const int yr[3][2] = {{0, tl[1]}, {tl[1], th[1]}, {th[1], n[1]}};
    const int zr[3][2] = {{0, tl[2]}, {tl[2], th[2]}, {th[2], n[2]}};
    for (int zi = 0; zi < 3; ++zi)
        for (int yi = 0; yi < 3; ++yi) {
            const int y0 = yr[yi][0], y1 = yr[yi][1], z0 = zr[zi][0], z1 = zr[zi][1];
            if (y1 <= y0 || z1 <= z0) continue;
            const int axes = (yi != 1 ? 2 : 0) | (zi != 1 ? 4 : 0);
            if (mode == 1) {
                Box m{0, n[0], y0, y1, z0, z1, 32, 0, 0, 0, 0};
                fit_columns(m, 32, axes | 1);
                items.push_back({scene, m, axes | 1});
            } else {
                Box m{tl[0], th[0], y0, y1, z0, z1, 64, 0, 0, 0, 0};
                fit_columns(m, axes ? 32 : 64, axes);
                items.push_back({scene, m, axes});
                Box e{0, n[0], y0, y1, z0, z1, lprx, 2, 0, 0, 0};
                e.tstride = n[0] - Wx;
                items.push_back({scene, e, axes | 1});
            }
        }
}

// Groups items by variant (one launch each) into `plan`, with row groups and
// z slices: per launch the slice length (`slice` planes, or 0: chosen from
// candidates) minimises the launch-round estimate sum of the tiles' sweeps /
// CUs + the longest sweep, a sweep being the slice's planes + 3.
// var_first / var_total: each variant's offset in the workgroup table and
// tile count; returns the algorithmic bytes of one pass.
uint64_t group_items(const std::vector<PassItem> &items, const std::vector<const Geo *> &geo,
                     int slice, int cus, std::vector<PassItem> &plan, int var_first[8],
                     int var_total[8])
{
    plan.clear();
    uint64_t first = 0, bytes = 0;
    for (int v = 0; v < 8; ++v) {
        std::vector<PassItem> vi;
        for (const auto &it : items)
            if (it.axes == v) vi.push_back(it);
        int depth = 0;  // the deepest box of the variant
        for (auto &it : vi) {
            const int rows = twostep_batch_rows_per_tile(v, it.box.lpr);
            it.box.gy = (it.box.y1 - it.box.y0 + rows - 1) / rows;
            depth = std::max(depth, it.box.z1 - it.box.z0);
        }
        static const int cand[] = {8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 1 << 30};
        int pick = slice;
        if (pick <= 0) {
            double best = -1.0;
            for (int S : cand) {
                double sum = 0.0;
                for (const auto &it : vi) {
                    const int nz = it.box.z1 - it.box.z0, gz = (nz + S - 1) / S;
                    sum += (double)it.box.gx * it.box.gy * (nz + 3.0 * gz);
                }
                const double cost = sum / cus + (double)std::min(S, depth) + 3.0;
                if (best < 0.0 || cost < best) {
                    best = cost;
                    pick = S;
                }
                if (S >= depth) break;
            }
        }
        var_first[v] = (int)first;
        uint64_t tiles = 0;
        for (auto &it : vi) {
            const int nz = it.box.z1 - it.box.z0;
            it.box.gz = std::min(nz, (nz + pick - 1) / std::max(pick, 1));
            it.box.first = (int)tiles;
            tiles += (uint64_t)it.box.gx * it.box.gy * it.box.gz;
            bytes += twostep_box_bytes(*geo[it.scene], it.box);
            plan.push_back(it);
        }
        var_total[v] = (int)tiles;
        first += tiles;
    }
    return bytes;
}

void plan_boxes(wf_ctx *c)
{
    const Geo &g = c->g;
    const int n[3] = {g.nx, g.ny, g.nz};
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        if (g.T[a] == 0) {
            lo[a] = 0;
            hi[a] = n[a];
        } else if (a == 0) {
            // halo x-1 outside the layer; rows start on a quad, and on a
            // 128-byte line when that moves at most one more quad into the
            // (slower, narrow-row) frame kernels
            const int q = (g.T[a] + 1 + 3) / 4 * 4, l = (g.T[a] + 1 + 31) / 32 * 32;
            lo[a] = l - q <= 4 ? l : q;
            hi[a] = n[a] - g.T[a];
        } else {
            lo[a] = g.T[a] + 1;
            hi[a] = n[a] - g.T[a];
        }
        if (hi[a] < lo[a]) hi[a] = lo[a];
    }
    const int zc = c->zchunk, fz = c->fzchunk;
    const int fr = fused_rows_per_tile(false), frx = fused_rows_per_tile(true);
    auto slices = [&](long tiles, int planes) {
        return c->zchunk_env ? (planes + zc - 1) / zc : fused_slices(tiles, planes, c->cus);
    };
    Box in{lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], 64, 0, 0, 0, 0};
    const bool empty = hi[0] <= lo[0] || hi[1] <= lo[1] || hi[2] <= lo[2];
    if (empty) {
        in = Box{0, 0, 0, 0, 0, 0, 64, 0, 0, 0, 0};
        lo[0] = hi[0] = 0;
        lo[1] = hi[1] = 0;
        lo[2] = hi[2] = 0;
    } else {
        in.gx = (hi[0] - lo[0] + 255) / 256;
        in.gy = (hi[1] - lo[1] + fr - 1) / fr;
        in.gz = slices((long)in.gx * in.gy, hi[2] - lo[2]);
    }
    c->interior = in;
    BoxList f{};
    if (empty) {
        add_box(f, split_box(0, n[0], 0, n[1], 0, n[2], fz));
    } else {
        add_box(f, split_box(0, n[0], 0, n[1], 0, lo[2], fz));
        add_box(f, split_box(0, n[0], 0, n[1], hi[2], n[2], fz));
        add_box(f, split_box(0, n[0], 0, lo[1], lo[2], hi[2], fz));
        add_box(f, split_box(0, n[0], hi[1], n[1], lo[2], hi[2], fz));
        add_box(f, split_box(0, lo[0], lo[1], hi[1], lo[2], hi[2], fz));
        add_box(f, split_box(hi[0], n[0], lo[1], hi[1], lo[2], hi[2], fz));
    }
    c->frame = f;
    // x fold: the x CPML cells of the interior's rows join the single pass
    // as two 256-cell edge columns (x-only CPML kernel, 8 waves), so that
    // no frame box has narrow rows (the x slabs, 24 cells wide at 1024^3
    // with T = 24, ran at about 3.5 TB/s against 5 for the y and z slabs).
    // Needs the x shell (whole blocks) within one edge column, the middle
    // columns' x-halo outside it, and no periodic x.
    c->xfold = false;
    const int W = 256, shx = g.sb[0] * 8;
    if (c->xfold_env && !empty && g.T[0] > 0 && !(g.periodic & 1u) && shx < W &&
        n[0] >= 2 * W) {
        Box e{0, n[0], lo[1], hi[1], lo[2], hi[2], 64, 2, 0, 0, 0};
        e.tstride = n[0] - W;
        const int fxr = fused_rows_per_tile_x();
        e.gy = (hi[1] - lo[1] + fxr - 1) / fxr;
        e.gz = slices(2l * e.gy, hi[2] - lo[2]);
        c->xedges = e;
        Box mid{W, n[0] - W, lo[1], hi[1], lo[2], hi[2], 64, 0, 0, 0, 0};
        mid.gx = (mid.x1 - mid.x0 + 255) / 256;
        mid.gy = (hi[1] - lo[1] + fr - 1) / fr;
        mid.gz = mid.gx > 0 ? slices((long)mid.gx * mid.gy, hi[2] - lo[2]) : 0;
        if (mid.x1 <= mid.x0) mid.gx = mid.gy = mid.gz = 0;
        c->xmid = mid;
        BoxList fx{};
        add_box(fx, split_box(0, n[0], 0, n[1], 0, lo[2], fz));
        add_box(fx, split_box(0, n[0], 0, n[1], hi[2], n[2], fz));
        add_box(fx, split_box(0, n[0], 0, lo[1], lo[2], hi[2], fz));
        add_box(fx, split_box(0, n[0], hi[1], n[1], lo[2], hi[2], fz));
        c->framex = fx;
        c->xfold = true;
    }
    // Two steps per pass: every region's tiles compute a margin two cells
    // below and one above their box on each axis (x: a quad), which must be
    // outside the shell on the axes the region's variant treats as plain and
    // inside the domain or at its zero boundary.  Per axis the interior range
    // [lo, hi) keeps that margin outside the shell: x in [Wx, nx - Wx) with
    // Wx = T + 4 (the x shell ranges [0, Wx) and [nx - Wx, nx) run narrow rows
    // of Wx/4 + 2 lanes), y and z in [T + 2, n - max(T, 1) - 1) (the shell
    // ranges below and above take the rest).  No periodic axis.
    c->pair_geo = false;
    c->nregions = 0;
    const int Wx = g.T[0] + 4, lprx = Wx / 4 + 2;
    int tl[3], th[3];
    tl[0] = Wx;
    th[0] = n[0] - Wx;
    for (int a = 1; a < 3; ++a) {
        tl[a] = g.T[a] + 2;
        th[a] = n[a] - std::max(g.T[a], 1) - 1;
    }
    if (!g.periodic && lprx <= 32 && th[0] - tl[0] >= 8 && th[1] >= tl[1] && th[2] >= tl[2]) {
        const int yr[3][2] = {{0, tl[1]}, {tl[1], th[1]}, {th[1], n[1]}};
        const int zr[3][2] = {{0, tl[2]}, {tl[2], th[2]}, {th[2], n[2]}};
        auto tile = [&](Box b, int axes) {
            const int rows = twostep_rows_per_tile(axes, b.lpr);
            b.gy = (b.y1 - b.y0 + rows - 1) / rows;
            b.gz = c->zchunk_env ? (b.z1 - b.z0 + zc - 1) / zc
                                 : fused_slices((long)b.gx * b.gy, b.z1 - b.z0, c->cus, 3);
            b.gz = std::min(b.gz, b.z1 - b.z0);
            c->regions[c->nregions++] = wf_ctx::Region{b, axes};
        };
        for (int zi = 0; zi < 3; ++zi)
            for (int yi = 0; yi < 3; ++yi) {
                const int y0 = yr[yi][0], y1 = yr[yi][1], z0 = zr[zi][0], z1 = zr[zi][1];
                if (y1 <= y0 || z1 <= z0) continue;
                const int axes = (yi != 1 ? 2 : 0) | (zi != 1 ? 4 : 0);
                // the x interior range: rows of 64 lanes in the interior, 32
                // in the y and z shell ranges (taller tiles; measured at
                // 1024^3: interior 32 lanes 8.5-8.9 ms per step against
                // 6.7, shell ranges 64 lanes 2.40 ms per pass against 2.30)
                const int lpr = axes ? 32 : 64;
                Box m{tl[0], th[0], y0, y1, z0, z1, lpr, 0, 0, 0, 0};
                m.tstride = 4 * (lpr - 2);
                m.gx = ((th[0] - tl[0]) / 4 + lpr - 3) / (lpr - 2);
                tile(m, axes);
                // the two x shell ranges (56 cells wide at 1024^3 with 16-lane
                // rows: 7.15 ms per step against 6.74)
                Box e{0, n[0], y0, y1, z0, z1, lprx, 2, 0, 0, 0};
                e.tstride = n[0] - Wx;
                tile(e, axes | 1);
            }
        c->pair_geo = true;
    }
    // the rows schedule's items (pass_items, by width), grouped by variant
    {
        std::vector<PassItem> items;
        int mode = -1;
        if (const char *rp = std::getenv("WF_ROWS_PLAN")) mode = std::min(2, std::max(-1, std::atoi(rp)));
        pass_items(g, mode, 0, items);
        std::vector<const Geo *> geo{&g};
        c->rows_bytes = group_items(items, geo, 0, c->cus, c->rows_plan, c->rows_first, c->rows_total);
    }
    Box all{0, n[0], 0, n[1], 0, n[2], 64, 0, 0, 0, 0};
    all.gx = (n[0] + 255) / 256;
    all.gy = (n[1] + frx - 1) / frx;  // swept by the CPML or ghost variant
    all.gz = slices((long)all.gx * all.gy, n[2]);
    c->all = all;
    Box allg = all;
    {
        const int fg = fused_rows_per_tile_ghost();
        allg.gy = (n[1] + fg - 1) / fg;
        allg.gz = slices((long)allg.gx * allg.gy, n[2]);
    }
    c->allg = allg;
    BoxList w{};
    add_box(w, split_box(0, n[0], 0, n[1], 0, n[2], zc));
    c->whole = w;
}

// One full step: H phase, E phase, sources, probes (src/engine.rs:748-782).
//
// Fused variant: the frame's H kernel runs on a side stream alongside the
// interior kernel (both read state A and write disjoint cells of state B),
// then the frame's E kernel runs once both are done.  With ev (6 timing
// events) everything runs on the main stream, one kernel at a time, with an
// event pair around the fused (0, 1), H (2, 3) and E (4, 5) launches.
// Face hooks of one side: high faces run before the H phase (E^n and
// H^{n-1/2} in `es`), low faces between the phases (H^{n+1/2} in `hs`, E^n
// and the E-phase psi in `es`).
int run_hooks(wf_ctx *c, bool high, const Ptrs &es, const Ptrs &hs, hipStream_t st)
{
    const FaceHook *on[3];
    int n = 0;
    unsigned predicted = 0;
    for (int f = 0; f < 6; ++f)
        if (c->hook_on[f] && c->hook[f].mode != 1) predicted |= 1u << f;
    for (int f = high ? 1 : 0; f < 6; f += 2)
        if (c->hook_on[f]) {
            c->hook[f].predicted = predicted;
            on[n++] = &c->hook[f];
        }
    if (n) WF_HIP(launch_face_hooks(c->g, es, hs, on, n, c->d_step, st));
    return WF_OK;
}

// The interior box never touches a face: its launch needs no ghost reads.
StepLaunch plain(StepLaunch s)
{
    s.ghost = false;
    return s;
}

// The frame boxes are tiled with their own z-chunk.
StepLaunch framed(StepLaunch s, const wf_ctx *c)
{
    s.zchunk = c->fzchunk;
    return s;
}

// The hooks may run beside the interior kernel when every one of them reads
// and writes only frame cells and its own buffers: predicted or replayed
// domain faces whose normal axis has a shell of at least 4 cells (a hook
// reads depths 0..3).  Recording hooks may sit on interior cells.
bool hooks_beside_interior(const wf_ctx *c)
{
    bool any = false;
    for (int f = 0; f < 6; ++f) {
        if (!c->hook_on[f]) continue;
        if (c->hook[f].mode == 1 || c->g.T[f / 2] < 4) return false;
        any = true;
    }
    return any;
}

// The whole domain in one pass of the fused kernel (CPML shell included),
// when enabled (WF_PML_FUSED=1) and no learned face feeds ghost planes (its
// low faces need the H phase done before their E update).  Bit-identical
// to the frame kernels but slower on MI355X (1024^3 with CPML 24: 15.8 vs
// 12.8 ms per step): a wave with any shell lane takes the CPML path, whose
// psi loads are issued in the plane's arithmetic instead of a plane ahead,
// and the kernel needs 232 VGPRs; the frame kernels stay the default.
bool single_pass(const wf_ctx *c) { return c->fused && c->pml && c->pml_fused && !c->ghost_reads; }

// Learned faces on a domain without a CPML: one pass of the fused kernel
// over the whole domain (E beyond the high faces from their ghost planes,
// predicted before the pass), then the low faces' hooks (which read the
// new H) and the E kernel over the low faces' boundary cells, now with
// their ghost H.  No frame kernels.
bool learned_single_pass(const wf_ctx *c) { return c->fused && !c->pml && c->ghost_reads; }

// The folded x plan (plan_boxes) applies while no face hook is set: hooks
// on a domain with a CPML touch frame cells only.
bool folded(const wf_ctx *c)
{
    if (!c->xfold) return false;
    for (int f = 0; f < 6; ++f)
        if (c->hook_on[f]) return false;
    return true;
}

// The end of a step (src/engine.rs:770-782): sources, probes, then the
// device step counter.  slot: probes in the two-step kernel's box read the
// values it recorded (the first step of a pass).
int point_work(wf_ctx *c, const Ptrs &p, hipStream_t st, bool slot, bool sources = true)
{
    if (c->nsrc && sources) {
        hipLaunchKernelGGL(k_sources, dim3((c->nsrc_grp + 255) / 256), dim3(256), 0, st, p,
                           c->d_src, (int)c->nsrc_col, c->d_src_grp, (int)c->nsrc_grp,
                           c->d_srcval, c->d_step, c->src_base);
    }
    if (c->nprobe) {
        hipLaunchKernelGGL(k_probes, dim3((c->nprobe + 63) / 64), dim3(64), 0, st, p,
                           c->d_probe, (int)c->nprobe, (int)c->nfreq, c->d_cos, c->d_sin,
                           c->d_active, c->d_step, c->probe_base, c->d_sums, c->d_counts,
                           c->d_trace, slot ? c->d_pslot : nullptr);
    }
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, st, c->d_step);
    WF_HIP(hipGetLastError());
    return WF_OK;
}

int launch_step(wf_ctx *c, hipEvent_t *ev = nullptr)
{
    hipStream_t st = c->stream;
    if (learned_single_pass(c)) {
        StepLaunch s{c->g, c->p, c->alt, c->pml, c->zchunk, st};
        s.ghost = true;
        if (int rc = run_hooks(c, true, c->p, c->p, st)) return rc;
        if (ev) WF_HIP(hipEventRecord(ev[0], st));
        WF_HIP(c->lf(s, c->allg, false, st, false));
        if (ev) {
            WF_HIP(hipEventRecord(ev[1], st));
            WF_HIP(hipEventRecord(ev[2], st));
            WF_HIP(hipEventRecord(ev[3], st));
        }
        if (int rc = run_hooks(c, false, c->p, c->alt, st)) return rc;
        if (ev) WF_HIP(hipEventRecord(ev[4], st));
        WF_HIP(c->le(s, c->lofix, st));
        if (ev) WF_HIP(hipEventRecord(ev[5], st));
        std::swap(c->p, c->alt);
    } else if (single_pass(c)) {
        // hooks (recording only, no ghost planes): high faces before the
        // step (E^n, H^{n-1/2} in state A), low faces after it (H^{n+1/2} in
        // state B; E^n and the E-phase psi of step n still in state A)
        StepLaunch s{c->g, c->p, c->alt, c->pml, c->zchunk, st};
        if (int rc = run_hooks(c, true, c->p, c->p, st)) return rc;
        if (ev) WF_HIP(hipEventRecord(ev[0], st));
        WF_HIP(c->lf(s, c->all, true, st, false));
        if (ev) {
            WF_HIP(hipEventRecord(ev[1], st));
            for (int k = 2; k < 6; ++k) WF_HIP(hipEventRecord(ev[k], st));
        }
        if (int rc = run_hooks(c, false, c->p, c->alt, st)) return rc;
        std::swap(c->p, c->alt);
    } else if (c->fused && c->hook_overlap && !ev && c->side && c->frame.total > 0 &&
        c->interior.gx > 0 && hooks_beside_interior(c)) {
        // Face hooks and the frame's H phase on the side stream, beside the
        // interior kernel: the hooks touch only frame cells (the high faces
        // read state A, the low faces the frame's new H in state B), the
        // interior kernel reads state A and writes interior cells of B.
        // The frame's E phase waits for both.
        StepLaunch s{c->g, c->p, c->alt, c->pml, c->zchunk, st};
        s.ghost = c->ghost_reads;
        WF_HIP(hipEventRecord(c->fork, st));
        WF_HIP(hipStreamWaitEvent(c->side, c->fork, 0));
        if (int rc = run_hooks(c, true, c->p, c->p, c->side)) return rc;
        WF_HIP(c->lh(framed(s, c), c->frame, c->side));
        if (int rc = run_hooks(c, false, c->p, c->alt, c->side)) return rc;
        WF_HIP(hipEventRecord(c->join, c->side));
        WF_HIP(c->lf(plain(s), c->interior, false, st, false));
        WF_HIP(hipStreamWaitEvent(st, c->join, 0));
        WF_HIP(c->le(framed(s, c), c->frame, st));
        std::swap(c->p, c->alt);
    } else if (c->fused) {
        StepLaunch s{c->g, c->p, c->alt, c->pml, c->zchunk, st};
        s.ghost = c->ghost_reads;
        if (int rc = run_hooks(c, true, c->p, c->p, st)) return rc;
        const bool fold = folded(c);
        const BoxList &frame = fold ? c->framex : c->frame;
        // the single pass: the interior, or its middle and edge columns
        auto interior = [&](hipStream_t q) -> int {
            if (!fold) {
                WF_HIP(c->lf(plain(s), c->interior, false, q, false));
                return WF_OK;
            }
            if (c->xmid.gx > 0) WF_HIP(c->lf(plain(s), c->xmid, false, q, false));
            WF_HIP(c->lf(plain(s), c->xedges, true, q, true));
            return WF_OK;
        };
        const bool overlap = c->overlap && !ev && frame.total > 0 && c->interior.gx > 0 && c->side;
        if (overlap) {
            WF_HIP(hipEventRecord(c->fork, st));
            WF_HIP(hipStreamWaitEvent(c->side, c->fork, 0));
            WF_HIP(c->lh(framed(s, c), frame, c->side));
            WF_HIP(hipEventRecord(c->join, c->side));
            if (int rc = interior(st)) return rc;
            WF_HIP(hipStreamWaitEvent(st, c->join, 0));
        } else {
            if (ev) WF_HIP(hipEventRecord(ev[0], st));
            if (int rc = interior(st)) return rc;
            if (ev) WF_HIP(hipEventRecord(ev[1], st));
            if (ev) WF_HIP(hipEventRecord(ev[2], st));
            WF_HIP(c->lh(framed(s, c), frame, st));
            if (ev) WF_HIP(hipEventRecord(ev[3], st));
        }
        if (int rc = run_hooks(c, false, c->p, c->alt, st)) return rc;
        if (ev) WF_HIP(hipEventRecord(ev[4], st));
        WF_HIP(c->le(framed(s, c), frame, st));
        if (ev) WF_HIP(hipEventRecord(ev[5], st));
        std::swap(c->p, c->alt);
    } else {
        StepLaunch s{c->g, c->p, c->p, c->pml, c->zchunk, st};
        s.ghost = c->ghost_reads;
        if (int rc = run_hooks(c, true, c->p, c->p, st)) return rc;
        if (ev) {
            WF_HIP(hipEventRecord(ev[0], st));
            WF_HIP(hipEventRecord(ev[1], st));
            WF_HIP(hipEventRecord(ev[2], st));
        }
        WF_HIP(c->lh(s, c->whole, st));
        if (ev) WF_HIP(hipEventRecord(ev[3], st));
        if (int rc = run_hooks(c, false, c->p, c->p, st)) return rc;
        if (ev) WF_HIP(hipEventRecord(ev[4], st));
        WF_HIP(c->le(s, c->whole, st));
        if (ev) WF_HIP(hipEventRecord(ev[5], st));
    }
    if (int rc = point_work(c, c->p, st, false)) return rc;
    c->step += 1;
    return WF_OK;
}

// Two steps per pass apply to the fused variant without face hooks or the
// whole-domain CPML pass: on the regions when the geometry allows them, or
// in rows through the batched kernel (any geometry, not the contracted
// set); a batch runs them for any fused context without a periodic axis, so
// the point tables of the two-step kernels are built for all of those.
bool pair_tables(const wf_ctx *c) { return c->fused && c->pair_env && !c->g.periodic; }

// 0: single steps; 1: two-step passes on the regions; 2: in rows.
int pair_kind(const wf_ctx *c)
{
    if (!pair_tables(c) || c->pml_fused || c->sched == 2) return 0;
    for (int f = 0; f < 6; ++f)
        if (c->hook_on[f]) return 0;
    if (c->sched == 1 && !c->contract) return 2;
    return c->pair_geo ? 1 : 0;
}

bool pair_active(const wf_ctx *c) { return pair_kind(c) != 0; }

// The rows schedule's items (once) and their tables for both state
// parities, rewritten after the point tables or the buffers change.
int rows_tables(wf_ctx *c)
{
    if (pair_kind(c) != 2 || !c->rows_dirty) return WF_OK;
    if (!c->d_ritems || !c->d_rwg) {
        dfree(c->d_ritems);
        dfree(c->d_rwg);
        uint64_t tiles = 0;
        for (int v = 0; v < 8; ++v) tiles += (uint64_t)c->rows_total[v];
        std::vector<uint32_t> wg(tiles);
        for (size_t k = 0; k < c->rows_plan.size(); ++k) {
            const PassItem &it = c->rows_plan[k];
            const uint64_t cnt = (uint64_t)it.box.gx * it.box.gy * it.box.gz;
            for (uint64_t j = 0; j < cnt; ++j) wg[c->rows_first[it.axes] + it.box.first + j] = (uint32_t)k;
        }
        if (int rc = dmalloc(&c->d_rwg, std::max<uint64_t>(tiles, 1) * sizeof(uint32_t))) return rc;
        if (int rc = dmalloc(&c->d_ritems, 2 * c->rows_plan.size() * sizeof(TwoStepItem))) return rc;
        if (tiles) WF_HIP(hipMemcpy(c->d_rwg, wg.data(), tiles * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    const size_t m = c->rows_plan.size();
    std::vector<TwoStepItem> it(2 * m);
    for (size_t k = 0; k < 2 * m; ++k) {
        TwoStepItem &d = it[k];
        std::memset(&d, 0, sizeof d);
        d.g = c->g;
        d.a = k < m ? c->p : c->alt;
        d.b = k < m ? c->alt : c->p;
        d.box = c->rows_plan[k % m].box;
        d.pts.src = PointRows{c->src_pts.bits, c->src_pts.key, c->src_pts.ent, c->src_pts.n};
        d.pts.prb = PointRows{c->prb_pts.bits, c->prb_pts.key, c->prb_pts.ent, c->prb_pts.n};
        d.pts.values = c->d_srcval;
        d.pts.ncol = (int)c->nsrc_col;
        d.pts.base = c->src_base;
        d.pts.step = c->d_step;
        d.pts.slot = c->d_pslot;
    }
    // earlier passes may still read the tables
    WF_HIP(hipStreamSynchronize(c->stream));
    WF_HIP(hipMemcpy(c->d_ritems, it.data(), 2 * m * sizeof(TwoStepItem), hipMemcpyHostToDevice));
    c->rows_key = c->p.f[0];
    c->rows_dirty = false;
    return WF_OK;
}

// Two loop iterations in one pass (src/engine.rs:747-782 twice): the
// two-step kernels over every region, state A -> B, add the first step's
// sources to E(n+1) on the way and record its probes; then the first step's
// probes (from the slots) and step counter, and the second step's sources,
// probes and counter on B.  With ev (6 timing events): event pairs around
// the interior (0, 1), the x shell regions (2, 3) and the y and z shell
// regions (4, 5).
int launch_pair(wf_ctx *c, hipEvent_t *ev = nullptr)
{
    hipStream_t st = c->stream;
    if (pair_kind(c) == 2) {
        // rows: one batched launch per variant (event pair 0, 1 around all)
        if (int rc = rows_tables(c)) return rc;
        const size_t m = c->rows_plan.size();
        const TwoStepItem *items = c->d_ritems + (c->p.f[0] == c->rows_key ? 0 : m);
        if (ev) WF_HIP(hipEventRecord(ev[0], st));
        for (int v = 0; v < 8; ++v)
            WF_HIP(launch_batch_twostep(items, c->d_rwg + c->rows_first[v], c->rows_total[v], v, st));
        if (ev)
            for (int k = 1; k < 6; ++k) WF_HIP(hipEventRecord(ev[k], st));
        if (int rc = point_work(c, c->alt, st, true, false)) return rc;
        if (int rc = point_work(c, c->alt, st, false)) return rc;
        std::swap(c->p, c->alt);
        c->step += 2;
        return WF_OK;
    }
    TwoStepPoints pts;
    pts.src = PointRows{c->src_pts.bits, c->src_pts.key, c->src_pts.ent, c->src_pts.n};
    pts.prb = PointRows{c->prb_pts.bits, c->prb_pts.key, c->prb_pts.ent, c->prb_pts.n};
    pts.values = c->d_srcval;
    pts.ncol = (int)c->nsrc_col;
    pts.base = c->src_base;
    pts.step = c->d_step;
    pts.slot = c->d_pslot;
    StepLaunch s{c->g, c->p, c->alt, c->pml, c->fzchunk, st};
    // every region on one stream: spreading the shell regions' launches over
    // three more streams beside the interior's measured 7.02-7.28 ms per step
    // against 6.65-6.68 at 1024^3 (four runs each, one GPU)
    for (int grp = 0; grp < 3; ++grp) {
        if (ev) WF_HIP(hipEventRecord(ev[2 * grp], st));
        for (int k = 0; k < c->nregions; ++k) {
            const int ax = c->regions[k].axes;
            if ((grp == 0 && ax == 0) || (grp == 1 && (ax & 1)) || (grp == 2 && ax && !(ax & 1)))
                WF_HIP(c->lt(s, c->regions[k].box, pts, st, ax));
        }
        if (ev) WF_HIP(hipEventRecord(ev[2 * grp + 1], st));
    }
    if (int rc = point_work(c, c->alt, st, true, false)) return rc;
    if (int rc = point_work(c, c->alt, st, false)) return rc;
    std::swap(c->p, c->alt);
    c->step += 2;
    return WF_OK;
}

// n loop iterations, two per pass where they apply.
int launch_steps(wf_ctx *c, uint64_t n)
{
    const bool pairs = pair_active(c);
    for (uint64_t s = 0; s < n;) {
        if (pairs && n - s >= 2) {
            if (int rc = launch_pair(c)) return rc;
            s += 2;
        } else {
            if (int rc = launch_step(c)) return rc;
            s += 1;
        }
    }
    return WF_OK;
}

// ---------------------------------------------------------------------------
// Field placement.  On MI355X the same kernels over the same 24 GiB state
// buffer run at speeds up to ~12% apart depending on the allocation, fixed
// for its life (1024^3 with CPML 24: the fused kernel reading one buffer
// and writing the other takes 8.2 to 9.6 ms, and A->B differs from B->A):
// it follows where the driver puts the pages physically, not the virtual
// layout.  For large grids wf_create therefore times a few steps on the
// zeroed state, then, for each further candidate, allocates one more state
// buffer while holding the current ones (so it lands elsewhere), times it in
// place of each state in turn, keeps the fastest arrangement and frees the
// buffer left over.  The state is zeroed again and the step counter reset
// afterwards.
// ---------------------------------------------------------------------------
constexpr int kPlaceMax = 8;       // extra state buffers tried
constexpr int kPlaceProbes = 17;   // step timings recorded (1 + 2 per extra buffer at most)

// Milliseconds per step over `timed` steps after `warm` untimed ones (both
// even, so the A/B states end where they started).  No sources, probes or
// hooks exist yet; the fields stay zero.
int probe_step_ms(wf_ctx *c, int warm, int timed, double *ms)
{
    hipEvent_t e0, e1;
    WF_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(WF_ERR_HIP, "placement probe: event creation failed");
    }
    int rc = launch_steps(c, (uint64_t)warm);
    if (rc == WF_OK && hipEventRecord(e0, c->stream) != hipSuccess) rc = fail(WF_ERR_HIP, "placement probe: event");
    if (rc == WF_OK) rc = launch_steps(c, (uint64_t)timed);
    float t = 0.f;
    if (rc == WF_OK && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&t, e0, e1) != hipSuccess))
        rc = fail(WF_ERR_HIP, "placement probe: timing failed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = (double)t / timed;
    return rc;
}

void point_fields(wf_ctx *c, uint64_t)
{
    // the batched kernel's item tables hold the state pointers: rewritten
    // before their next use
    c->rows_dirty = true;
    for (int i = 0; i < 6; ++i) {
        c->p.f[i] = c->d_fields + (uint64_t)i * c->fstride;
        c->alt.f[i] = c->d_fields2 ? c->d_fields2 + (uint64_t)i * c->fstride : nullptr;
    }
}

// Tries `extra` more state buffers (at most kPlaceMax) and keeps the fastest
// arrangement.  A buffer that cannot be allocated, or a search that has run
// for budget_ms, ends it (the current arrangement stays); HIP errors are
// returned.
// Back to the initial state after a probe: zero fields and psi, step 0.
int reset_state(wf_ctx *c)
{
    const uint64_t bytes = 6ull * c->fstride * sizeof(float);
    WF_HIP(hipMemsetAsync(c->d_fields, 0, bytes, c->stream));
    if (c->d_fields2) WF_HIP(hipMemsetAsync(c->d_fields2, 0, bytes, c->stream));
    if (c->d_psi) WF_HIP(hipMemsetAsync(c->d_psi, 0, c->psi_bytes, c->stream));
    if (c->d_psi2) WF_HIP(hipMemsetAsync(c->d_psi2, 0, c->psi_bytes, c->stream));
    WF_HIP(hipMemsetAsync(c->d_step, 0, sizeof(uint64_t), c->stream));
    WF_HIP(hipStreamSynchronize(c->stream));
    c->step = 0;
    return WF_OK;
}

int choose_placement(wf_ctx *c, int extra, uint64_t N, double budget_ms)
{
    const auto t_start = std::chrono::steady_clock::now();
    auto elapsed_ms = [&] {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    };
    const size_t bytes = 6 * c->fstride * sizeof(float);
    float **slot[2] = {&c->d_fields, &c->d_fields2};
    const int nslot = c->d_fields2 ? 2 : 1;
    c->place_n = 0;
    c->place_kept = 0;
    double best = 0.0;
    // three timed passes per arrangement (WF_PLACEMENT_PASSES): with one, the
    // kept arrangement's bench time came 0.02-0.35 ms per step above its
    // probe time at 1024^3, with three a steady 0.10-0.11 (four runs each)
    int timed = 6;
    if (const char *pp = std::getenv("WF_PLACEMENT_PASSES")) timed = 2 * std::max(1, std::atoi(pp));
    if (int rc = probe_step_ms(c, 2, timed, &best)) return rc;
    c->place_ms[c->place_n++] = best;
    for (int t = 0; t < extra && t < kPlaceMax && c->place_n + nslot <= kPlaceProbes && elapsed_ms() < budget_ms;
         ++t) {
        float *x = nullptr;
        if (hipMalloc(reinterpret_cast<void **>(&x), bytes) != hipSuccess) {
            (void)hipGetLastError();
            break;
        }
        int rc = WF_OK, win = -1;
        if (hipMemsetAsync(x, 0, bytes, c->stream) != hipSuccess)
            rc = fail(WF_ERR_HIP, "placement probe: hipMemset failed");
        for (int k = 0; k < nslot && rc == WF_OK; ++k) {
            float *held = *slot[k];
            *slot[k] = x;
            point_fields(c, N);
            double ms = 0.0;
            rc = probe_step_ms(c, 2, timed, &ms);
            *slot[k] = held;
            point_fields(c, N);
            if (rc) break;
            c->place_ms[c->place_n++] = ms;
            if (ms < best) {
                best = ms;
                win = k;
                c->place_kept = c->place_n - 1;
            }
        }
        (void)hipStreamSynchronize(c->stream);
        if (win >= 0) {
            (void)hipFree(*slot[win]);
            *slot[win] = x;
            point_fields(c, N);
        } else {
            (void)hipFree(x);
        }
        if (rc) return rc;
    }
    c->rows_dirty = true;
    return reset_state(c);
}

// Schedule of a small grid: the two-step regions (where the geometry allows
// them), the rows and single steps, each timed over 4 steps after 2 on the
// zeroed state; the fastest is kept.  Small scenes are launch- and
// latency-bound: the regions' 18 launches per pass cost most there.
int choose_schedule(wf_ctx *c)
{
    double best = -1.0;
    int pick = 2;
    for (int k = 0; k < 3; ++k) {
        c->sched = k;
        if (k == 0 && pair_kind(c) != 1) continue;
        double ms = 0.0;
        if (int rc = probe_step_ms(c, 2, 4, &ms)) return rc;
        c->sched_ms[k] = ms;
        if (best < 0.0 || ms < best) {
            best = ms;
            pick = k;
        }
    }
    c->sched = pick;
    return reset_state(c);
}

constexpr int kGraphSteps = 8;

// Captures kGraphSteps consecutive steps (an even number, so the fused
// variant's A/B state returns to the same buffers) into one executable graph.
// Step kernels read the step from the device counter, so the graph replays
// for any later step until the source or probe tables change.
int capture_graph(wf_ctx *c)
{
    drop_graph_exec(c);
    if (int rc = rows_tables(c)) return rc;  // no synchronous copy under capture
    const uint64_t step0 = c->step;
    const Ptrs p0 = c->p, a0 = c->alt;
    hipGraph_t graph = nullptr;
    WF_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_steps(c, kGraphSteps);
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    // capture only records: restore the host view of the state
    c->step = step0;
    c->p = p0;
    c->alt = a0;
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(WF_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    const hipError_t ie = hipGraphInstantiate(&c->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
        c->gexec = nullptr;
        return fail(WF_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ie));
    }
    c->gstate = c->p.f[0];
    return WF_OK;
}

int ensure_stage(wf_ctx *c, uint64_t bytes)
{
    if (c->stage_bytes >= bytes) return WF_OK;
    dfree(c->d_stage);
    c->stage_bytes = 0;
    if (int rc = dmalloc(&c->d_stage, bytes)) return rc;
    c->stage_bytes = bytes;
    return WF_OK;
}

// Block planes per staging chunk: at most ~256 MiB.
uint64_t planes_per_chunk(const wf_ctx *c)
{
    const uint64_t plane_bytes = (uint64_t)c->g.nb[0] * c->g.nb[1] * 512 * 24;
    uint64_t n = (256ull << 20) / plane_bytes;
    if (n < 1) n = 1;
    if (n > (uint64_t)c->g.nb[2]) n = c->g.nb[2];
    return n;
}

// Packs the whole current state into `out` (the reference's snapshot byte
// order, src/engine.rs:622-640) on the main stream, without waiting.
int pack_snapshot(wf_ctx *c, float *out)
{
    const uint64_t plane_vox = (uint64_t)c->g.nb[0] * c->g.nb[1] * 512;
    uint64_t ppc = (1ull << 30) / plane_vox;  // keeps one launch under 2^22 workgroups
    if (ppc < 1) ppc = 1;
    for (uint64_t bz = 0; bz < (uint64_t)c->g.nb[2]; bz += ppc) {
        const uint64_t np = std::min<uint64_t>(ppc, c->g.nb[2] - bz);
        const uint64_t nvox = np * plane_vox;
        hipLaunchKernelGGL(k_pack, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, c->stream,
                           c->p, c->g, (int)bz, nvox, out + bz * plane_vox * 6);
        WF_HIP(hipGetLastError());
    }
    return WF_OK;
}

}  // namespace

extern "C" {

int wf_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int wf_create(uint32_t nx, uint32_t ny, uint32_t nz, const uint8_t *material_ids,
              const int32_t *coeff_table, const uint32_t *pml_shell_blocks,
              const float *pml_profiles, uint32_t periodic_mask, int device, wf_ctx **out)
{
    clear_error();
    if (!out) return fail(WF_ERR_INVALID, "wf_create: out is null");
    *out = nullptr;
    if (!nx || !ny || !nz || nx % 8 || ny % 8 || nz % 8)
        return fail(WF_ERR_INVALID,
                    "grid dimensions must be positive multiples of 8 (got %ux%ux%u)", nx, ny, nz);
    if (nx > (1u << 20) || ny > (1u << 20) || nz > (1u << 20))
        return fail(WF_ERR_INVALID, "grid axis longer than 2^20 voxels");
    if (!coeff_table) return fail(WF_ERR_INVALID, "wf_create: coeff_table is null");
    if (pml_shell_blocks && !pml_profiles)
        return fail(WF_ERR_INVALID, "wf_create: PML shell given without profiles");
    if (periodic_mask > 7u) return fail(WF_ERR_INVALID, "wf_create: periodic_mask > 7");
    if (fault_injected("wf_create")) return fail(WF_ERR_HIP, "fault injected: wf_create");

    const uint32_t n[3] = {nx, ny, nz};
    Geo g{};
    g.nx = (int)nx;
    g.ny = (int)ny;
    g.nz = (int)nz;
    for (int a = 0; a < 3; ++a) g.nb[a] = (int)(n[a] / 8);
    g.periodic = periodic_mask;
    if (pml_shell_blocks) {
        const float *pp = pml_profiles;
        for (int a = 0; a < 3; ++a) {
            const uint32_t sb = pml_shell_blocks[a];
            if (2ull * sb > (uint64_t)g.nb[a])
                return fail(WF_ERR_INVALID,
                            "PML shell of %u blocks on axis %d exceeds half the axis (%d blocks)",
                            sb, a, g.nb[a]);
            g.sb[a] = (int)sb;
            g.T[a] = (int)(8 * sb);
            for (uint32_t i = 0; i < n[a]; ++i) {
                const bool in_layer = i < 8 * sb || i >= n[a] - 8 * sb;
                if (!in_layer && (pp[i] != 1.0f || pp[n[a] + i] != 0.0f || pp[2 * n[a] + i] != 1.0f))
                    return fail(WF_ERR_INVALID,
                                "PML profile on axis %d is not identity at %u, outside the "
                                "graded layer of %u voxels",
                                a, i, 8 * sb);
            }
            pp += 3 * n[a];
        }
    }

    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0)
        return fail(WF_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= ndev)
        return fail(WF_ERR_INVALID, "device %d out of range (%d devices)", device, ndev);
    WF_HIP(hipSetDevice(device));

    wf_ctx *c = new wf_ctx();
    c->device = device;
    c->g = g;
    c->pml = pml_shell_blocks != nullptr;
    auto bail = [&](int rc) {
        wf_destroy(c);
        return rc;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->join, hipEventDisableTiming) != hipSuccess)
        return bail(fail(WF_ERR_HIP, "stream / event creation failed"));

    if (const char *kv = std::getenv("WF_KERNEL")) {
        if (std::strcmp(kv, "split") == 0) c->fused = false;
        else if (std::strcmp(kv, "fused") == 0) c->fused = true;
        else return bail(fail(WF_ERR_INVALID, "WF_KERNEL must be 'fused' or 'split' (got '%s')", kv));
    }
    const uint64_t N = (uint64_t)nx * ny * nz;
    // the six fields of a state lie N + 1024 floats apart (WF_FIELD_PAD bytes,
    // whole 256 B, instead of 4096): with N a power of two the fields would
    // start at power-of-two distances; measured at 1024^3 the two-step
    // interior kernel takes 9.23-9.25 ms per launch against 9.34-9.51
    // unpadded (four runs each, one GPU), the step 0.1-0.3% less
    uint64_t pad = 4096;
    if (const char *fp = std::getenv("WF_FIELD_PAD")) pad = (uint64_t)std::max(0L, std::atol(fp));
    c->fstride = N + pad / 256 * 64;
    const uint64_t sbytes = 6 * c->fstride * sizeof(float);
    if (int rc = dmalloc(&c->d_fields, sbytes)) return bail(rc);
    if (c->fused) {
        if (int rc = dmalloc(&c->d_fields2, sbytes)) return bail(rc);
        if (hipMemsetAsync(c->d_fields2, 0, sbytes, c->stream) != hipSuccess)
            return bail(fail(WF_ERR_HIP, "hipMemset(fields) failed"));
    }
    if (int rc = dmalloc(&c->d_id, N)) return bail(rc);
    if (int rc = dmalloc(&c->d_coef, 256 * sizeof(float4))) return bail(rc);
    if (int rc = dmalloc(&c->d_step, sizeof(uint64_t))) return bail(rc);
    if (hipMemsetAsync(c->d_step, 0, sizeof(uint64_t), c->stream) != hipSuccess)
        return bail(fail(WF_ERR_HIP, "hipMemset(step) failed"));
    point_fields(c, N);
    c->p.id = c->alt.id = c->d_id;
    c->p.coef = c->alt.coef = c->d_coef;

    if (hipMemsetAsync(c->d_fields, 0, sbytes, c->stream) != hipSuccess)
        return bail(fail(WF_ERR_HIP, "hipMemset(fields) failed"));
    if (material_ids) {
        if (hipMemcpyAsync(c->d_id, material_ids, N, hipMemcpyHostToDevice, c->stream) != hipSuccess)
            return bail(fail(WF_ERR_HIP, "upload of material ids failed"));
    } else if (hipMemsetAsync(c->d_id, 0, N, c->stream) != hipSuccess) {
        return bail(fail(WF_ERR_HIP, "hipMemset(ids) failed"));
    }
    // Dequantize once: FixedQ16_16::to_f32 (src/layout.rs:123-125).
    std::vector<float4> coef(256);
    for (int i = 0; i < 256; ++i)
        coef[i] = make_float4(wf_q16_to_f32(coeff_table[4 * i + 0]), wf_q16_to_f32(coeff_table[4 * i + 1]),
                              wf_q16_to_f32(coeff_table[4 * i + 2]), wf_q16_to_f32(coeff_table[4 * i + 3]));
    if (hipMemcpyAsync(c->d_coef, coef.data(), 256 * sizeof(float4), hipMemcpyHostToDevice,
                       c->stream) != hipSuccess)
        return bail(fail(WF_ERR_HIP, "upload of coefficient table failed"));

    // Profiles (identity when PML is off) and psi slabs.
    {
        std::vector<float> prof(3 * (nx + ny + nz));
        if (pml_profiles) {
            std::memcpy(prof.data(), pml_profiles, prof.size() * sizeof(float));
        } else {
            float *pp = prof.data();
            for (int a = 0; a < 3; ++a) {
                for (uint32_t i = 0; i < n[a]; ++i) {
                    pp[i] = 1.0f;
                    pp[n[a] + i] = 0.0f;
                    pp[2 * n[a] + i] = 1.0f;
                }
                pp += 3 * n[a];
            }
        }
        if (int rc = dmalloc(&c->d_prof, prof.size() * sizeof(float))) return bail(rc);
        if (hipMemcpy(c->d_prof, prof.data(), prof.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(WF_ERR_HIP, "upload of PML profiles failed"));
        float *dp = c->d_prof;
        for (int a = 0; a < 3; ++a) {
            for (int j = 0; j < 3; ++j) c->p.prof[a][j] = c->alt.prof[a][j] = dp + j * n[a];
            dp += 3 * n[a];
        }
    }
    uint64_t slab[3] = {2ull * g.T[0] * ny * nz, 2ull * g.T[1] * nx * nz, 2ull * g.T[2] * nx * ny};
    const uint64_t psi_total = 4 * (slab[0] + slab[1] + slab[2]);
    c->psi_bytes = psi_total * sizeof(float);
    if (psi_total) {
        if (int rc = dmalloc(&c->d_psi, psi_total * sizeof(float))) return bail(rc);
        if (hipMemsetAsync(c->d_psi, 0, psi_total * sizeof(float), c->stream) != hipSuccess)
            return bail(fail(WF_ERR_HIP, "hipMemset(psi) failed"));
        if (c->fused) {
            if (int rc = dmalloc(&c->d_psi2, psi_total * sizeof(float))) return bail(rc);
            if (hipMemsetAsync(c->d_psi2, 0, psi_total * sizeof(float), c->stream) != hipSuccess)
                return bail(fail(WF_ERR_HIP, "hipMemset(psi) failed"));
        }
    }
    {
        float *pp = c->d_psi, *pp2 = c->d_psi2;
        for (int a = 0; a < 3; ++a)
            for (int t = 0; t < 4; ++t) {
                c->p.psi[a][t] = slab[a] ? pp : nullptr;
                c->alt.psi[a][t] = (slab[a] && pp2) ? pp2 : nullptr;
                pp += slab[a];
                if (pp2) pp2 += slab[a];
            }
    }
    // Shell block list for wf_read_psi.
    if (c->pml) {
        std::vector<uint64_t> list;
        for (int bz = 0; bz < g.nb[2]; ++bz)
            for (int by = 0; by < g.nb[1]; ++by)
                for (int bx = 0; bx < g.nb[0]; ++bx) {
                    const int bc[3] = {bx, by, bz};
                    bool in = false;
                    for (int a = 0; a < 3; ++a)
                        if (g.sb[a] > 0 && (bc[a] < g.sb[a] || bc[a] >= g.nb[a] - g.sb[a])) in = true;
                    if (in) list.push_back((uint64_t)bx | ((uint64_t)by << 20) | ((uint64_t)bz << 40));
                }
        c->nshell_blocks = list.size();
        if (!list.empty()) {
            if (int rc = dmalloc(&c->d_shell_list, list.size() * sizeof(uint64_t))) return bail(rc);
            if (hipMemcpy(c->d_shell_list, list.data(), list.size() * sizeof(uint64_t),
                          hipMemcpyHostToDevice) != hipSuccess)
                return bail(fail(WF_ERR_HIP, "upload of shell list failed"));
        }
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess)
        return bail(fail(WF_ERR_HIP, "initialisation failed"));
    // Planes per workgroup of the split kernels over the whole domain (split
    // variant) and of the low-face fix-up boxes: z-marches reuse E(z+1) /
    // H(z-1) from registers, but shorter ones give more workgroups; 8 planes
    // is the fastest at 1024^3 (H + E 15.41 ms per step, against 15.47 at 16,
    // 15.78 at 32, 15.64 at 4 and 17.96 at 1, same allocation), halved for
    // small grids until >= 4096 workgroups.  WF_ZCHUNK overrides.
    {
        const long tiles_xy = (long)((nx + 255) / 256) * (long)((ny + 6) / 7);
        int z = 8;
        while (z > 2 && tiles_xy * (long)((nz + z - 1) / z) < 4096) z /= 2;
        c->zchunk = z;
    }
    if (const char *zc = std::getenv("WF_ZCHUNK")) {
        const int v = std::atoi(zc);
        if (v > 0) {
            c->zchunk = v;
            c->zchunk_env = true;
        }
    }
    {
        int cu = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cu > 0)
            c->cus = cu;
    }
    // The frame kernels (thin slabs around the interior) run fastest with
    // one plane per workgroup: at 1024^3 with CPML 24 the frame's H + E take
    // 3.51 ms per step at 1 plane, 3.56 at 4, 3.60 at 8 and 3.75 at 32.
    if (const char *fz = std::getenv("WF_FRAME_ZCHUNK")) {
        const int v = std::atoi(fz);
        if (v > 0) c->fzchunk = v;
    }
    if (const char *ov = std::getenv("WF_OVERLAP")) c->overlap = std::atoi(ov) != 0;
    if (const char *ho = std::getenv("WF_HOOK_OVERLAP")) c->hook_overlap = std::atoi(ho) != 0;
    if (const char *pf = std::getenv("WF_PML_FUSED")) c->pml_fused = std::atoi(pf) != 0;
    if (const char *xf = std::getenv("WF_XFOLD")) c->xfold_env = std::atoi(xf) != 0;
    if (const char *gr = std::getenv("WF_GRAPH")) c->use_graph = std::atoi(gr) != 0;
    if (const char *fc = std::getenv("WF_FP_CONTRACT"))
        if (std::atoi(fc) != 0) {
            c->contract = true;
            c->lh = fma::launch_h;
            c->le = fma::launch_e;
            c->lf = fma::launch_fused;
            c->lt = fma::launch_twostep;
        }
    if (const char *ts = std::getenv("WF_TWOSTEP")) c->pair_env = std::atoi(ts) != 0;
    plan_boxes(c);
    // Field placement (see choose_placement): by default eight extra state
    // buffers tried, one held at a time beside the current ones, when one
    // state's fields take at least 8 GiB (2 + 6 steps per arrangement, 17
    // arrangements: about 1 s at 1024^3), no new buffer after 4 s.
    // WF_PLACEMENT=n tries n (0 = no probe, at most 8); WF_PLACEMENT_MIN_MIB
    // moves the size threshold.
    {
        int tries = 8;
        uint64_t min_bytes = 8ull << 30;
        if (const char *pl = std::getenv("WF_PLACEMENT")) tries = std::atoi(pl);
        if (const char *mb = std::getenv("WF_PLACEMENT_MIN_MIB"))
            min_bytes = (uint64_t)std::max(0L, std::atol(mb)) << 20;
        if (tries > 0 && 6 * N * sizeof(float) >= min_bytes)
            if (int rc = choose_placement(c, std::min(tries, kPlaceMax), N, 4000.0)) return bail(rc);
    }
    // Schedule (choose_schedule) for grids under 2^25 cells, else the
    // regions; WF_SCHEDULE=0 (regions), 1 (rows) or 2 (single steps) sets it.
    if (const char *sc = std::getenv("WF_SCHEDULE")) {
        c->sched = std::min(2, std::max(0, std::atoi(sc)));
    } else if (pair_tables(c) && !c->pml_fused && !c->contract && N < (1ull << 25)) {
        if (int rc = choose_schedule(c)) return bail(rc);
    }
    *out = c;
    return WF_OK;
}

int wf_placement(const wf_ctx *c, double *ms, int cap, int *n, int *kept)
{
    clear_error();
    if (!c || !n || !kept || (cap > 0 && !ms)) return fail(WF_ERR_INVALID, "wf_placement: null pointer");
    *n = c->place_n;
    *kept = c->place_kept;
    for (int i = 0; i < c->place_n && i < cap; ++i) ms[i] = c->place_ms[i];
    return WF_OK;
}

void wf_destroy(wf_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->snap_stream) (void)hipStreamSynchronize(c->snap_stream);
    dfree(c->d_snap);
    if (c->snap_packed) (void)hipEventDestroy(c->snap_packed);
    if (c->snap_done) (void)hipEventDestroy(c->snap_done);
    if (c->snap_stream) (void)hipStreamDestroy(c->snap_stream);
    free_sources(c);
    free_probes(c);
    dfree(c->d_fields);
    dfree(c->d_fields2);
    dfree(c->d_psi2);
    dfree(c->d_ritems);
    dfree(c->d_rwg);
    dfree(c->d_step);
    dfree(c->d_id);
    dfree(c->d_coef);
    dfree(c->d_prof);
    dfree(c->d_psi);
    dfree(c->d_shell_list);
    dfree(c->d_stage);
    for (int f = 0; f < 6; ++f) {
        dfree(c->d_ghost[f]);
        dfree(c->d_mem[f]);
        dfree(c->d_samples[f]);
        dfree(c->d_net[f]);
    }
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->join) (void)hipEventDestroy(c->join);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int wf_set_source_columns(wf_ctx *c, uint32_t n, const wf_source_site *sites,
                          const uint32_t *cols, uint32_t ncol, uint64_t nsteps,
                          const float *values)
{
    clear_error();
    if (!c) return fail(WF_ERR_INVALID, "wf_set_source_columns: null context");
    WF_HIP(hipSetDevice(c->device));
    WF_HIP(hipStreamSynchronize(c->stream));
    free_sources(c);
    if (n == 0) return WF_OK;
    if (!sites || (!values && nsteps && ncol)) return fail(WF_ERR_INVALID, "wf_set_sources: null pointer");
    if (n > 0x7fffffffu) return fail(WF_ERR_INVALID, "wf_set_sources: more than 2^31 - 1 sources");
    if (ncol == 0) return fail(WF_ERR_INVALID, "wf_set_source_columns: no value columns");
    if (nsteps > (~0ull >> 2) / ncol) return fail(WF_ERR_INVALID, "wf_set_sources: value table too large");
    std::vector<DevSource> ds(n);
    for (uint32_t s = 0; s < n; ++s) {
        const wf_source_site &q = sites[s];
        if (q.x >= (uint32_t)c->g.nx || q.y >= (uint32_t)c->g.ny || q.z >= (uint32_t)c->g.nz)
            return fail(WF_ERR_INVALID, "source %u at (%u, %u, %u) is outside the domain", s, q.x, q.y, q.z);
        if (q.component < 0 || q.component > 2)
            return fail(WF_ERR_INVALID, "source %u: component %d is not Ex/Ey/Ez", s, q.component);
        const uint32_t col = cols ? cols[s] : s;
        if (col >= ncol)
            return fail(WF_ERR_INVALID, "source %u: column %u of %u", s, col, ncol);
        ds[s].off = ((uint64_t)q.z * c->g.ny + q.y) * c->g.nx + q.x;
        ds[s].comp = q.component;
        ds[s].hard = q.hard ? 1 : 0;
        ds[s].col = col;
        ds[s].pad = 0;
    }
    // group by (component, voxel), keeping list order inside a group
    std::stable_sort(ds.begin(), ds.end(), [](const DevSource &a, const DevSource &b) {
        return a.comp != b.comp ? a.comp < b.comp : a.off < b.off;
    });
    std::vector<uint32_t> grp;
    for (uint32_t s = 0; s < n; ++s)
        if (s == 0 || ds[s].comp != ds[s - 1].comp || ds[s].off != ds[s - 1].off) grp.push_back(s);
    grp.push_back(n);
    if (int rc = dmalloc(&c->d_src, n * sizeof(DevSource))) return rc;
    if (int rc = dmalloc(&c->d_src_grp, grp.size() * sizeof(uint32_t))) return rc;
    if (int rc = dmalloc(&c->d_srcval, std::max<uint64_t>(nsteps * ncol, 1) * sizeof(float))) return rc;
    WF_HIP(hipMemcpy(c->d_src, ds.data(), n * sizeof(DevSource), hipMemcpyHostToDevice));
    WF_HIP(hipMemcpy(c->d_src_grp, grp.data(), grp.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    c->nsrc_grp = (uint32_t)(grp.size() - 1);
    if (nsteps) WF_HIP(hipMemcpy(c->d_srcval, values, nsteps * ncol * sizeof(float), hipMemcpyHostToDevice));
    if (pair_tables(c)) {
        // the two-step kernel's table: by row, list order kept per voxel
        std::vector<std::pair<uint64_t, PointEnt>> v(n);
        for (uint32_t s = 0; s < n; ++s) {
            PointEnt e{};
            e.x = (uint32_t)(ds[s].off % (uint64_t)c->g.nx);
            e.comp = (uint16_t)ds[s].comp;
            e.hard = (uint16_t)ds[s].hard;
            e.col = ds[s].col;
            v[s] = {ds[s].off / (uint64_t)c->g.nx, e};
        }
        if (int rc = upload_points(c, v, c->src_pts)) return rc;
    }
    c->nsrc = n;
    c->nsrc_col = ncol;
    c->src_base = c->step;
    c->src_steps = nsteps;
    return WF_OK;
}

int wf_set_sources(wf_ctx *c, uint32_t n, const wf_source_site *sites, uint64_t nsteps,
                   const float *values)
{
    return wf_set_source_columns(c, n, sites, nullptr, n, nsteps, values);
}

int wf_set_probes(wf_ctx *c, uint32_t n, const wf_probe_site *sites, uint64_t nsteps,
                  const double *cos_tab, const double *sin_tab, const uint8_t *active, int trace)
{
    clear_error();
    if (!c) return fail(WF_ERR_INVALID, "wf_set_probes: null context");
    WF_HIP(hipSetDevice(c->device));
    WF_HIP(hipStreamSynchronize(c->stream));
    free_probes(c);
    if (n == 0) return WF_OK;
    if (!sites || (nsteps && !active)) return fail(WF_ERR_INVALID, "wf_set_probes: null pointer");
    std::vector<DevProbe> dp(n);
    uint32_t nf = 0;
    for (uint32_t p = 0; p < n; ++p) {
        const wf_probe_site &q = sites[p];
        if (q.x >= (uint32_t)c->g.nx || q.y >= (uint32_t)c->g.ny || q.z >= (uint32_t)c->g.nz)
            return fail(WF_ERR_INVALID, "probe position (%u, %u, %u) is outside the %dx%dx%d domain",
                        q.x, q.y, q.z, c->g.nx, c->g.ny, c->g.nz);
        if (q.component < 0 || q.component > 5)
            return fail(WF_ERR_INVALID, "probe %u: component %d out of range", p, q.component);
        dp[p].off = ((uint64_t)q.z * c->g.ny + q.y) * c->g.nx + q.x;
        dp[p].comp = q.component;
        dp[p].nfreq = (int)q.nfreq;
        dp[p].foff = (int)nf;
        nf += q.nfreq;
    }
    if (nf && nsteps && (!cos_tab || !sin_tab))
        return fail(WF_ERR_INVALID, "wf_set_probes: null twiddle table");
    if (int rc = dmalloc(&c->d_probe, n * sizeof(DevProbe))) return rc;
    if (int rc = dmalloc(&c->d_cos, nsteps * nf * sizeof(double))) return rc;
    if (int rc = dmalloc(&c->d_sin, nsteps * nf * sizeof(double))) return rc;
    if (int rc = dmalloc(&c->d_sums, 2ull * (nf ? nf : 1) * sizeof(double))) return rc;
    if (int rc = dmalloc(&c->d_active, nsteps * n)) return rc;
    if (int rc = dmalloc(&c->d_counts, n * sizeof(unsigned long long))) return rc;
    if (trace) {
        if (int rc = dmalloc(&c->d_trace, nsteps * n * sizeof(float))) return rc;
    }
    WF_HIP(hipMemcpy(c->d_probe, dp.data(), n * sizeof(DevProbe), hipMemcpyHostToDevice));
    if (nsteps && nf) {
        WF_HIP(hipMemcpy(c->d_cos, cos_tab, nsteps * nf * sizeof(double), hipMemcpyHostToDevice));
        WF_HIP(hipMemcpy(c->d_sin, sin_tab, nsteps * nf * sizeof(double), hipMemcpyHostToDevice));
    }
    if (nsteps) WF_HIP(hipMemcpy(c->d_active, active, nsteps * n, hipMemcpyHostToDevice));
    WF_HIP(hipMemset(c->d_sums, 0, 2ull * (nf ? nf : 1) * sizeof(double)));
    WF_HIP(hipMemset(c->d_counts, 0, n * sizeof(unsigned long long)));
    if (pair_tables(c)) {
        // the two-step regions own every cell: each probe's value of a
        // pass's first step is recorded into its slot by the tile owning it
        std::vector<std::pair<uint64_t, PointEnt>> v(n);
        for (uint32_t p = 0; p < n; ++p) {
            const wf_probe_site &q = sites[p];
            PointEnt e{};
            e.x = q.x;
            e.comp = (uint16_t)q.component;
            e.col = p;
            v[p] = {(uint64_t)q.z * c->g.ny + q.y, e};
        }
        if (int rc = upload_points(c, v, c->prb_pts)) return rc;
        if (int rc = dmalloc(&c->d_pslot, n * sizeof(float))) return rc;
        WF_HIP(hipMemset(c->d_pslot, 0, n * sizeof(float)));
    }
    c->nprobe = n;
    c->nfreq = nf;
    c->probe_base = c->step;
    c->probe_steps = nsteps;
    return WF_OK;
}

int wf_step(wf_ctx *c, uint64_t nsteps)
{
    clear_error();
    if (!c) return fail(WF_ERR_INVALID, "wf_step: null context");
    if (fault_injected("wf_step")) return fail(WF_ERR_HIP, "fault injected: wf_step");
    if (c->nsrc && c->step + nsteps > c->src_base + c->src_steps)
        return fail(WF_ERR_STATE, "wf_step: source table covers %llu steps, %llu requested past step %llu",
                    (unsigned long long)c->src_steps, (unsigned long long)nsteps,
                    (unsigned long long)(c->step - c->src_base));
    if (c->nprobe && c->step + nsteps > c->probe_base + c->probe_steps)
        return fail(WF_ERR_STATE, "wf_step: probe table covers %llu steps", (unsigned long long)c->probe_steps);
    WF_HIP(hipSetDevice(c->device));
    uint64_t left = nsteps;
    if (c->use_graph && left >= (uint64_t)kGraphSteps) {
        // the graph was captured with a given A/B buffer assignment
        if (c->gexec && c->p.f[0] != c->gstate) {
            if (int rc = launch_step(c)) return rc;
            --left;
        }
        if (!c->gexec && left >= (uint64_t)kGraphSteps)
            if (int rc = capture_graph(c)) return rc;
        while (c->gexec && left >= (uint64_t)kGraphSteps) {
            WF_HIP(hipGraphLaunch(c->gexec, c->stream));
            c->step += kGraphSteps;
            left -= kGraphSteps;
        }
    }
    return launch_steps(c, left);
}

int wf_sync(wf_ctx *c)
{
    if (!c) return fail(WF_ERR_INVALID, "wf_sync: null context");
    WF_HIP(hipSetDevice(c->device));
    WF_HIP(hipStreamSynchronize(c->stream));
    return WF_OK;
}

uint64_t wf_step_count(const wf_ctx *c) { return c ? c->step : 0; }

uint64_t wf_field_bytes(const wf_ctx *c)
{
    return c ? (uint64_t)c->g.nx * c->g.ny * c->g.nz * 24ull : 0;
}

int wf_read_fields(wf_ctx *c, void *dst)
{
    clear_error();
    if (!c || !dst) return fail(WF_ERR_INVALID, "wf_read_fields: null pointer");
    if (fault_injected("wf_read_fields")) return fail(WF_ERR_HIP, "fault injected: wf_read_fields");
    WF_HIP(hipSetDevice(c->device));
    const uint64_t ppc = planes_per_chunk(c);
    const uint64_t plane_vox = (uint64_t)c->g.nb[0] * c->g.nb[1] * 512;
    if (int rc = ensure_stage(c, ppc * plane_vox * 24)) return rc;
    for (uint64_t bz = 0; bz < (uint64_t)c->g.nb[2]; bz += ppc) {
        const uint64_t np = std::min<uint64_t>(ppc, c->g.nb[2] - bz);
        const uint64_t nvox = np * plane_vox;
        hipLaunchKernelGGL(k_pack, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, c->stream,
                           c->p, c->g, (int)bz, nvox, c->d_stage);
        WF_HIP(hipGetLastError());
        WF_HIP(hipMemcpyAsync(static_cast<char *>(dst) + bz * plane_vox * 24, c->d_stage, nvox * 24,
                              hipMemcpyDeviceToHost, c->stream));
        WF_HIP(hipStreamSynchronize(c->stream));
    }
    return WF_OK;
}

int wf_snapshot_async(wf_ctx *c, void *dst)
{
    clear_error();
    if (!c || !dst) return fail(WF_ERR_INVALID, "wf_snapshot_async: null pointer");
    WF_HIP(hipSetDevice(c->device));
    const uint64_t bytes = (uint64_t)c->g.nx * c->g.ny * c->g.nz * 24;
    if (!c->d_snap) {
        if (int rc = dmalloc(&c->d_snap, bytes)) return rc;
        if (hipStreamCreateWithFlags(&c->snap_stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&c->snap_packed, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c->snap_done, hipEventDisableTiming) != hipSuccess)
            return fail(WF_ERR_HIP, "wf_snapshot_async: stream or event creation failed");
    }
    // the previous snapshot must have left the device buffer before it is packed again
    if (c->snap_pending) WF_HIP(hipStreamWaitEvent(c->stream, c->snap_done, 0));
    if (int rc = pack_snapshot(c, c->d_snap)) return rc;
    WF_HIP(hipEventRecord(c->snap_packed, c->stream));
    WF_HIP(hipStreamWaitEvent(c->snap_stream, c->snap_packed, 0));
    WF_HIP(hipMemcpyAsync(dst, c->d_snap, bytes, hipMemcpyDeviceToHost, c->snap_stream));
    WF_HIP(hipEventRecord(c->snap_done, c->snap_stream));
    c->snap_pending = true;
    return WF_OK;
}

int wf_wait(wf_ctx *c)
{
    clear_error();
    if (!c) return fail(WF_ERR_INVALID, "wf_wait: null context");
    if (!c->snap_pending) return WF_OK;
    WF_HIP(hipSetDevice(c->device));
    WF_HIP(hipEventSynchronize(c->snap_done));
    c->snap_pending = false;
    return WF_OK;
}

int wf_run(wf_ctx *c, uint64_t num_steps, uint64_t snapshot_every, const char *output_path,
           double *seconds)
{
    clear_error();
    if (!c || !output_path) return fail(WF_ERR_INVALID, "wf_run: null pointer");
    if (snapshot_every == 0) return fail(WF_ERR_INVALID, "wf_run: snapshot_every must be at least 1");
    if (c->nsrc && c->step + num_steps > c->src_base + c->src_steps)
        return fail(WF_ERR_STATE, "wf_run: source table covers %llu steps, %llu requested",
                    (unsigned long long)c->src_steps, (unsigned long long)num_steps);
    if (c->nprobe && c->step + num_steps > c->probe_base + c->probe_steps)
        return fail(WF_ERR_STATE, "wf_run: probe table covers %llu steps", (unsigned long long)c->probe_steps);
    WF_HIP(hipSetDevice(c->device));
    const auto t0 = std::chrono::steady_clock::now();
    const uint64_t bytes = wf_field_bytes(c);
    SnapshotWriter writer;
    int rc = writer.open(output_path, bytes, c->device);
    float *snap[2] = {nullptr, nullptr};
    hipEvent_t packed[2] = {nullptr, nullptr};
    if (!rc) rc = dmalloc(&snap[0], bytes);
    if (!rc) rc = dmalloc(&snap[1], bytes);
    for (int i = 0; i < 2 && !rc; ++i)
        if (hipEventCreateWithFlags(&packed[i], hipEventDisableTiming) != hipSuccess)
            rc = fail(WF_ERR_HIP, "wf_run: event creation failed");
    // Same schedule as the reference loop (src/engine.rs:747-799): after the
    // step with loop index k, a snapshot when k % snapshot_every == 0.  Steps
    // between snapshots go to wf_step in one call (graph replays).
    uint64_t k = 0;
    int slot = 0;
    while (!rc && k < num_steps) {
        const uint64_t next = (k + snapshot_every - 1) / snapshot_every * snapshot_every;
        const uint64_t upto = std::min(next, num_steps - 1);
        rc = wf_step(c, upto - k + 1);
        if (rc) break;
        k = upto + 1;
        if (upto % snapshot_every) continue;
        // the only point where the loop waits on storage: the slot's
        // previous snapshot has not been written yet
        rc = writer.wait_free(slot);
        if (!rc) rc = pack_snapshot(c, snap[slot]);
        if (!rc && hipEventRecord(packed[slot], c->stream) != hipSuccess)
            rc = fail(WF_ERR_HIP, "wf_run: event record failed");
        if (!rc) rc = writer.submit(slot, snap[slot], packed[slot]);
        slot ^= 1;
    }
    std::string msg = rc ? wf_last_error() : "";
    const int wrc = writer.finish();  // drains every queued snapshot
    if (hipStreamSynchronize(c->stream) != hipSuccess && !rc && !wrc) {
        rc = WF_ERR_HIP;
        msg = "wf_run: stream synchronize failed";
    }
    for (auto &e : packed)
        if (e) (void)hipEventDestroy(e);
    dfree(snap[0]);
    dfree(snap[1]);
    if (rc) return fail(rc, "%s", msg.c_str());
    if (wrc) return wrc;
    if (seconds)
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return WF_OK;
}

int wf_write_fields(wf_ctx *c, const void *src)
{
    clear_error();
    if (!c || !src) return fail(WF_ERR_INVALID, "wf_write_fields: null pointer");
    WF_HIP(hipSetDevice(c->device));
    const uint64_t ppc = planes_per_chunk(c);
    const uint64_t plane_vox = (uint64_t)c->g.nb[0] * c->g.nb[1] * 512;
    if (int rc = ensure_stage(c, ppc * plane_vox * 24)) return rc;
    for (uint64_t bz = 0; bz < (uint64_t)c->g.nb[2]; bz += ppc) {
        const uint64_t np = std::min<uint64_t>(ppc, c->g.nb[2] - bz);
        const uint64_t nvox = np * plane_vox;
        WF_HIP(hipMemcpyAsync(c->d_stage, static_cast<const char *>(src) + bz * plane_vox * 24, nvox * 24,
                              hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_unpack, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, c->stream,
                           c->p, c->g, (int)bz, nvox, c->d_stage);
        WF_HIP(hipGetLastError());
        WF_HIP(hipStreamSynchronize(c->stream));
    }
    return WF_OK;
}

uint64_t wf_psi_shell_blocks(const wf_ctx *c) { return c ? c->nshell_blocks : 0; }

int wf_read_psi(wf_ctx *c, float *dst)
{
    clear_error();
    if (!c || !dst) return fail(WF_ERR_INVALID, "wf_read_psi: null pointer");
    if (!c->nshell_blocks) return WF_OK;
    WF_HIP(hipSetDevice(c->device));
    const uint64_t per = 12ull * 512 * sizeof(float);
    uint64_t chunk = (256ull << 20) / per;
    if (int rc = ensure_stage(c, std::min<uint64_t>(chunk, c->nshell_blocks) * per)) return rc;
    for (uint64_t b = 0; b < c->nshell_blocks; b += chunk) {
        const uint64_t nb = std::min<uint64_t>(chunk, c->nshell_blocks - b);
        hipLaunchKernelGGL(k_pack_psi, dim3((unsigned)((nb * 512 + 255) / 256)), dim3(256), 0,
                           c->stream, c->p, c->g, c->d_shell_list + b, nb, c->d_stage);
        WF_HIP(hipGetLastError());
        WF_HIP(hipMemcpyAsync(reinterpret_cast<char *>(dst) + b * per, c->d_stage, nb * per,
                              hipMemcpyDeviceToHost, c->stream));
        WF_HIP(hipStreamSynchronize(c->stream));
    }
    return WF_OK;
}

int wf_read_probe_sums(wf_ctx *c, double *sum_cos, double *sum_sin, uint64_t *n)
{
    clear_error();
    if (!c) return fail(WF_ERR_INVALID, "wf_read_probe_sums: null context");
    if (!c->nprobe) return WF_OK;
    WF_HIP(hipSetDevice(c->device));
    WF_HIP(hipStreamSynchronize(c->stream));
    if (c->nfreq) {
        WF_HIP(hipMemcpy(sum_cos, c->d_sums, c->nfreq * sizeof(double), hipMemcpyDeviceToHost));
        WF_HIP(hipMemcpy(sum_sin, c->d_sums + c->nfreq, c->nfreq * sizeof(double), hipMemcpyDeviceToHost));
    }
    static_assert(sizeof(unsigned long long) == sizeof(uint64_t), "count width");
    WF_HIP(hipMemcpy(n, c->d_counts, c->nprobe * sizeof(uint64_t), hipMemcpyDeviceToHost));
    return WF_OK;
}

int wf_read_trace(wf_ctx *c, float *trace, uint64_t max_steps)
{
    clear_error();
    if (!c || !trace) return fail(WF_ERR_INVALID, "wf_read_trace: null pointer");
    if (!c->d_trace) return fail(WF_ERR_STATE, "wf_read_trace: probes were set without a trace");
    WF_HIP(hipSetDevice(c->device));
    WF_HIP(hipStreamSynchronize(c->stream));
    uint64_t k = c->step - c->probe_base;
    if (k > max_steps) k = max_steps;
    if (k) WF_HIP(hipMemcpy(trace, c->d_trace, k * c->nprobe * sizeof(float), hipMemcpyDeviceToHost));
    return WF_OK;
}

int wf_time_kernels(wf_ctx *c, uint64_t nsteps, double *fused_ms, double *h_ms, double *e_ms)
{
    clear_error();
    if (!c || !fused_ms || !h_ms || !e_ms)
        return fail(WF_ERR_INVALID, "wf_time_kernels: null pointer");
    if (c->nsrc && c->step + nsteps > c->src_base + c->src_steps)
        return fail(WF_ERR_STATE, "wf_time_kernels: source table too short");
    if (c->nprobe && c->step + nsteps > c->probe_base + c->probe_steps)
        return fail(WF_ERR_STATE, "wf_time_kernels: probe table too short");
    WF_HIP(hipSetDevice(c->device));
    hipEvent_t ev[6];
    for (auto &e : ev) WF_HIP(hipEventCreate(&e));
    double acc[3] = {0.0, 0.0, 0.0};
    int rc = WF_OK;
    // two-step passes: the two-step kernel per launch, and the frame of the
    // pass's first and of its second step
    const bool pairs = pair_active(c) && nsteps >= 2;
    const uint64_t n = pairs ? nsteps / 2 : nsteps;
    for (uint64_t s = 0; s < n && rc == WF_OK; ++s) {
        rc = pairs ? launch_pair(c, ev) : launch_step(c, ev);
        if (rc) break;
        WF_HIP(hipEventSynchronize(ev[5]));
        for (int k = 0; k < 3; ++k) {
            float ms = 0.f;
            WF_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
            acc[k] += ms;
        }
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    const double d = n ? (double)n : 1.0;
    // the split variant launches no fused kernel (its event pair is adjacent)
    *fused_ms = c->fused ? acc[0] / d : 0.0;
    *h_ms = acc[1] / d;
    *e_ms = acc[2] / d;
    return rc;
}

int wf_hbm_copy(uint64_t bytes, int device, int iters, double *ms_per_copy)
{
    clear_error();
    if (!ms_per_copy || bytes < 16 || iters < 1)
        return fail(WF_ERR_INVALID, "wf_hbm_copy: bad arguments");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return fail(WF_ERR_HIP, "wf_hbm_copy: no device %d", device);
    WF_HIP(hipSetDevice(device));
    const uint64_t n4 = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    if (int rc = dmalloc(&a, n4 * 16)) return rc;
    if (int rc = dmalloc(&b, n4 * 16)) {
        dfree(a);
        return rc;
    }
    hipStream_t s;
    hipEvent_t e0, e1;
    int rc = WF_OK;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        dfree(a);
        dfree(b);
        return fail(WF_ERR_HIP, "wf_hbm_copy: stream");
    }
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipMemsetAsync(a, 0, n4 * 16, s);
    if ((n4 + 255) / 256 > 0x7fffffffull) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
        dfree(a);
        dfree(b);
        return fail(WF_ERR_INVALID, "wf_hbm_copy: copy too large");
    }
    const dim3 grid((unsigned)((n4 + 255) / 256));
    hipLaunchKernelGGL(k_copy, grid, dim3(256), 0, s, a, b, n4);  // warm
    (void)hipEventRecord(e0, s);
    for (int i = 0; i < iters; ++i) hipLaunchKernelGGL(k_copy, grid, dim3(256), 0, s, a, b, n4);
    (void)hipEventRecord(e1, s);
    if (hipEventSynchronize(e1) != hipSuccess) rc = fail(WF_ERR_HIP, "wf_hbm_copy: sync failed");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *ms_per_copy = (double)ms / iters;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    dfree(a);
    dfree(b);
    return rc;
}

const char *wf_kernel_variant(const wf_ctx *c)
{
    if (!c) return "";
    if (!c->fused) return "split";
    const int k = pair_kind(c);
    return k == 1 ? "twostep" : (k == 2 ? "twostep-rows" : "fused");
}

int wf_schedule(const wf_ctx *c, int *schedule, double *probe_ms)
{
    if (!c || !schedule) return fail(WF_ERR_INVALID, "wf_schedule: null pointer");
    *schedule = c->sched;
    if (probe_ms)
        for (int k = 0; k < 3; ++k) probe_ms[k] = c->sched_ms[k];
    return WF_OK;
}

int wf_kernel_bytes(const wf_ctx *c, uint64_t *fused_bytes, uint64_t *h_bytes,
                    uint64_t *e_bytes)
{
    if (!c || !fused_bytes || !h_bytes || !e_bytes)
        return fail(WF_ERR_INVALID, "wf_kernel_bytes: null pointer");
    const uint64_t nx = c->g.nx, ny = c->g.ny, nz = c->g.nz;
    // non-identity psi terms per phase: 2 per graded axis per cell
    const uint64_t terms = 2ull * (2ull * c->g.T[0] * ny * nz + 2ull * c->g.T[1] * nx * nz +
                                   2ull * c->g.T[2] * nx * ny);
    if (pair_kind(c) == 2) {
        // rows: one pass over every cell (the batched kernels, one event pair)
        *fused_bytes = c->rows_bytes;
        *h_bytes = *e_bytes = 0;
        return WF_OK;
    }
    if (pair_active(c)) {
        // two-step tiles over the whole domain: 49 B per cell for two steps
        // and 16 B per psi term (2 per phase of each graded axis of a cell,
        // read once and written once per pass); the interior, the x shell
        // regions, the y and z shell regions
        const Geo &g = c->g;
        uint64_t grp[3] = {0, 0, 0};
        for (int k = 0; k < c->nregions; ++k) {
            const Box &b = c->regions[k].box;
            const int ax = c->regions[k].axes;
            // the x shell regions: two columns of nx - tstride cells
            const uint64_t wx = (ax & 1) ? 2ull * (uint64_t)(b.x1 - b.tstride) : (uint64_t)(b.x1 - b.x0);
            const uint64_t wy = b.y1 - b.y0, wz = b.z1 - b.z0;
            auto gr = [&](int a, int lo, int hi, int n) {
                uint64_t k2 = 0;
                for (int q = lo; q < hi; ++q) k2 += (g.T[a] > 0 && (q < g.T[a] || q >= n - g.T[a])) ? 1 : 0;
                return k2;
            };
            const uint64_t gxc = (ax & 1) ? 2ull * g.T[0] : 0;
            const uint64_t psi = 2 * (gxc * wy * wz + gr(1, b.y0, b.y1, g.ny) * wx * wz +
                                      gr(2, b.z0, b.z1, g.nz) * wx * wy);
            const int gi = ax == 0 ? 0 : ((ax & 1) ? 1 : 2);
            grp[gi] += 49ull * wx * wy * wz + 16ull * psi;
        }
        *fused_bytes = grp[0];
        *h_bytes = grp[1];
        *e_bytes = grp[2];
        return WF_OK;
    }
    if (single_pass(c)) {
        // one pass: 49 B per cell, and per psi term of each phase 4 B read
        // and 4 B written
        *fused_bytes = 49ull * nx * ny * nz + 16ull * terms;
        *h_bytes = *e_bytes = 0;
        return WF_OK;
    }
    if (learned_single_pass(c)) {
        // one pass, then the E kernel over the low faces' boundary cells
        // (37 B per cell)
        uint64_t cells = 0;
        for (int k = 0; k < c->lofix.n; ++k) {
            const Box &b = c->lofix.b[k];
            cells += (uint64_t)(b.x1 - b.x0) * (b.y1 - b.y0) * (b.z1 - b.z0);
        }
        *fused_bytes = 49ull * nx * ny * nz;
        *h_bytes = 0;
        *e_bytes = 37ull * cells;
        return WF_OK;
    }
    const bool fold = c->fused && folded(c);
    const BoxList &bl = c->fused ? (fold ? c->framex : c->frame) : c->whole;
    uint64_t cells = 0;
    for (int k = 0; k < bl.n; ++k) {
        const Box &b = bl.b[k];
        cells += (uint64_t)(b.x1 - b.x0) * (b.y1 - b.y0) * (b.z1 - b.z0);
    }
    const Box &in = c->interior;
    uint64_t fb = 0, fterms = 0;
    if (fold) {
        // the middle and edge columns: 49 B per cell, and 16 B (read and
        // write, both phases) per x psi term of the edges' graded cells
        const Box &m = c->xmid, &e = c->xedges;
        if (m.x1 > m.x0) fb += 49ull * (uint64_t)(m.x1 - m.x0) * (m.y1 - m.y0) * (m.z1 - m.z0);
        fb += 49ull * 2 * 256 * (uint64_t)(e.y1 - e.y0) * (e.z1 - e.z0);
        fterms = 2ull * 2ull * c->g.T[0] * (uint64_t)(in.y1 - in.y0) * (in.z1 - in.z0);
        fb += 16ull * fterms;
    } else if (c->fused) {
        fb = 49ull * (uint64_t)(in.x1 - in.x0) * (in.y1 - in.y0) * (in.z1 - in.z0);
    }
    *fused_bytes = fb;
    *h_bytes = 37ull * cells + 8ull * (terms - fterms);
    *e_bytes = 37ull * cells + 8ull * (terms - fterms);
    return WF_OK;
}

int wf_time_steps(wf_ctx *c, uint64_t nsteps, double *total_ms)
{
    clear_error();
    if (!c || !total_ms) return fail(WF_ERR_INVALID, "wf_time_steps: null pointer");
    WF_HIP(hipSetDevice(c->device));
    hipEvent_t e0, e1;
    WF_HIP(hipEventCreate(&e0));
    WF_HIP(hipEventCreate(&e1));
    WF_HIP(hipEventRecord(e0, c->stream));
    int rc = wf_step(c, nsteps);
    WF_HIP(hipEventRecord(e1, c->stream));
    WF_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    WF_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *total_ms = ms;
    return rc;
}

// ===========================================================================
// Learned boundary faces
// ===========================================================================
}  // extern "C"

namespace {

struct FaceGeom {
    int a, hi, na, nf, ns;
};

FaceGeom face_geom(const wf_ctx *c, int f)
{
    const int n[3] = {c->g.nx, c->g.ny, c->g.nz};
    FaceGeom fg;
    fg.a = f / 2;
    fg.hi = f % 2;
    const int t1 = (fg.a + 1) % 3, t2 = (fg.a + 2) % 3;
    fg.na = n[fg.a];
    fg.nf = n[std::min(t1, t2)];
    fg.ns = n[std::max(t1, t2)];
    return fg;
}

void clear_face(wf_ctx *c, int f)
{
    dfree(c->d_ghost[f]);
    dfree(c->d_mem[f]);
    dfree(c->d_samples[f]);
    dfree(c->d_net[f]);
    dfree(c->d_pool[f]);
    c->hook_on[f] = false;
    c->hook[f] = FaceHook{};
    for (int q = 0; q < 2; ++q) c->p.ghost[f][q] = c->alt.ghost[f][q] = nullptr;
}

void update_ghost_reads(wf_ctx *c)
{
    c->ghost_reads = false;
    for (int f = 0; f < 6; ++f)
        if (c->hook_on[f] && c->hook[f].mode != 1) c->ghost_reads = true;
    // boundary cells of the low faces with ghost planes (x: the first quad
    // of every row; cells on two such faces are computed twice, with the
    // same result, since the E kernel reads only H and the old E)
    const Geo &g = c->g;
    const int zc = c->zchunk;
    BoxList lf{};
    auto ghosted = [&](int f) { return c->hook_on[f] && c->hook[f].mode != 1; };
    if (ghosted(4)) add_box(lf, split_box(0, g.nx, 0, g.ny, 0, 1, zc));
    if (ghosted(2)) add_box(lf, split_box(0, g.nx, 0, 1, 0, g.nz, zc));
    if (ghosted(0)) add_box(lf, split_box(0, 4, 0, g.ny, 0, g.nz, zc));
    c->lofix = lf;
}

}  // namespace

extern "C" {

int wf_set_fp_contract(wf_ctx *c, int on)
{
    clear_error();
    if (!c) return fail(WF_ERR_INVALID, "wf_set_fp_contract: null context");
    WF_HIP(hipSetDevice(c->device));
    WF_HIP(hipStreamSynchronize(c->stream));
    drop_graph_exec(c);
    c->contract = on != 0;
    c->lh = on ? fma::launch_h : launch_h;
    c->le = on ? fma::launch_e : launch_e;
    c->lf = on ? fma::launch_fused : launch_fused;
    c->lt = on ? fma::launch_twostep : launch_twostep;
    return WF_OK;
}

int wf_clear_faces(wf_ctx *c)
{
    clear_error();
    if (!c) return fail(WF_ERR_INVALID, "wf_clear_faces: null context");
    WF_HIP(hipSetDevice(c->device));
    WF_HIP(hipStreamSynchronize(c->stream));
    drop_graph_exec(c);
    for (int f = 0; f < 6; ++f) clear_face(c, f);
    update_ghost_reads(c);
    return WF_OK;
}

}  // extern "C"

namespace {
// Energy credit (zeroed) and flux partials of face f when its ghosts are
// limited; none otherwise.
int set_pool(wf_ctx *c, int f, int passive)
{
    if (!passive) {
        c->hook[f].credit = c->hook[f].parts = nullptr;
        return WF_OK;
    }
    const size_t bytes = (16 + (size_t)kPartsFloats) * sizeof(float);
    if (!c->d_pool[f])
        if (int rc = dmalloc(&c->d_pool[f], bytes)) return rc;
    if (hipMemset(c->d_pool[f], 0, bytes) != hipSuccess) return fail(WF_ERR_HIP, "credit clear failed");
    c->hook[f].credit = c->d_pool[f];
    c->hook[f].parts = c->d_pool[f] + 16;
    return WF_OK;
}
}  // namespace

extern "C" {

int wf_set_face_limits(wf_ctx *c, int passive, float retain, float gamma)
{
    clear_error();
    if (!c) return fail(WF_ERR_INVALID, "wf_set_face_limits: null context");
    if (!(gamma >= 0.f) || !std::isfinite(gamma))
        return fail(WF_ERR_INVALID, "wf_set_face_limits: gamma must be finite and >= 0");
    if (!(retain >= 0.f && retain < 1.f))
        return fail(WF_ERR_INVALID, "wf_set_face_limits: retain must lie in [0, 1)");
    WF_HIP(hipSetDevice(c->device));
    WF_HIP(hipStreamSynchronize(c->stream));
    drop_graph_exec(c);
    c->face_passive = passive != 0;
    c->face_gamma = gamma;
    c->face_retain = retain;
    for (int f = 0; f < 6; ++f)
        if (c->hook_on[f] && c->hook[f].mode == 0) {
            if (int rc = set_pool(c, f, c->face_passive)) return rc;
            c->hook[f].passive = c->face_passive;
            c->hook[f].gamma = c->face_gamma;
            c->hook[f].retain = c->face_retain;
        }
    return WF_OK;
}

int wf_set_replay_limits(wf_ctx *c, int passive, float retain, float gamma)
{
    clear_error();
    if (!c) return fail(WF_ERR_INVALID, "wf_set_replay_limits: null context");
    if (!(gamma >= 0.f) || !std::isfinite(gamma))
        return fail(WF_ERR_INVALID, "wf_set_replay_limits: gamma must be finite and >= 0");
    if (!(retain >= 0.f && retain < 1.f))
        return fail(WF_ERR_INVALID, "wf_set_replay_limits: retain must lie in [0, 1)");
    WF_HIP(hipSetDevice(c->device));
    WF_HIP(hipStreamSynchronize(c->stream));
    drop_graph_exec(c);
    for (int f = 0; f < 6; ++f)
        if (c->hook_on[f] && c->hook[f].mode == 2) {
            if (passive && !c->d_mem[f]) {
                const uint64_t bytes = 2ull * c->hook[f].nfast * c->hook[f].nslow * kMemStride * sizeof(float);
                if (int rc = dmalloc(&c->d_mem[f], bytes)) return rc;
                WF_HIP(hipMemset(c->d_mem[f], 0, bytes));
                c->hook[f].mem = c->d_mem[f];
            }
            if (int rc = set_pool(c, f, passive != 0)) return rc;
            c->hook[f].passive = passive != 0;
            c->hook[f].gamma = gamma;
            c->hook[f].retain = retain;
        }
    return WF_OK;
}

int wf_set_learned_faces(wf_ctx *c, uint32_t face_mask, const float *params, int head)
{
    clear_error();
    if (!c) return fail(WF_ERR_INVALID, "wf_set_learned_faces: null context");
    if (!valid_head(head))
        return fail(WF_ERR_INVALID, "wf_set_learned_faces: unknown head %d", head);
    if (face_mask >= 64u) return fail(WF_ERR_INVALID, "wf_set_learned_faces: face mask has bits above 5");
    if (face_mask && !params) return fail(WF_ERR_INVALID, "wf_set_learned_faces: null parameters");
    for (int f = 0; f < 6; ++f) {
        if (!(face_mask >> f & 1u)) continue;
        const FaceGeom fg = face_geom(c, f);
        if (c->g.periodic >> fg.a & 1u)
            return fail(WF_ERR_INVALID, "face %d lies on a periodic axis", f);
        if (fg.na < 8) return fail(WF_ERR_INVALID, "face %d: axis shorter than 8 cells", f);
        if (c->pml && c->g.T[fg.a] < 4)
            return fail(WF_ERR_INVALID, "face %d: in a domain with a CPML its axis needs a shell "
                        "of at least 4 cells (the frame kernels then hold the face)", f);
    }
    WF_HIP(hipSetDevice(c->device));
    WF_HIP(hipStreamSynchronize(c->stream));
    drop_graph_exec(c);
    float *d_par = nullptr;
    if (face_mask)
        if (int rc = dmalloc(&d_par, (size_t)kMlpParams * sizeof(float))) return rc;
    int rc = WF_OK;
    for (int f = 0; f < 6 && !rc; ++f) {
        if (!(face_mask >> f & 1u)) continue;
        clear_face(c, f);
        const FaceGeom fg = face_geom(c, f);
        const uint64_t npos = (uint64_t)fg.nf * fg.ns;
        rc = dmalloc(&c->d_ghost[f], 2 * npos * sizeof(float));
        if (!rc) rc = dmalloc(&c->d_mem[f], 2 * npos * kMemStride * sizeof(float));
        if (!rc) rc = dmalloc(&c->d_net[f], sizeof(MlpPacked));
        if (rc) break;
        if (hipMemset(c->d_ghost[f], 0, 2 * npos * sizeof(float)) != hipSuccess ||
            hipMemset(c->d_mem[f], 0, 2 * npos * kMemStride * sizeof(float)) != hipSuccess ||
            hipMemcpy(d_par, params + (size_t)f * kMlpParams, kMlpParams * sizeof(float),
                      hipMemcpyHostToDevice) != hipSuccess ||
            pack_mlp(d_par, head, c->d_net[f], c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
            rc = fail(WF_ERR_HIP, "wf_set_learned_faces: upload failed");
            break;
        }
        FaceHook &h = c->hook[f];
        h.axis = fg.a;
        h.hi = fg.hi;
        h.b0 = fg.hi ? fg.na - 1 : 0;
        h.fast0 = 0;
        h.nfast = fg.nf;
        h.slow0 = 0;
        h.nslow = fg.ns;
        h.na = fg.na;
        h.nf = fg.nf;
        h.ns = fg.ns;
        h.mode = 0;
        h.head = head;
        if ((rc = set_pool(c, f, c->face_passive))) break;
        h.passive = c->face_passive;
        h.gamma = c->face_gamma;
        h.retain = c->face_retain;
        h.ghost[0] = c->d_ghost[f];
        h.ghost[1] = c->d_ghost[f] + npos;
        h.mem = c->d_mem[f];
        h.net = c->d_net[f];
        for (int q = 0; q < 2; ++q) c->p.ghost[f][q] = c->alt.ghost[f][q] = h.ghost[q];
        c->hook_on[f] = true;
    }
    dfree(d_par);
    update_ghost_reads(c);
    return rc;
}

int wf_record_face_samples(wf_ctx *c, uint32_t face_mask, const wf_face_region *regions,
                           uint64_t every, uint64_t rows)
{
    clear_error();
    if (!c || (face_mask && !regions)) return fail(WF_ERR_INVALID, "wf_record_face_samples: null pointer");
    if (face_mask >= 64u) return fail(WF_ERR_INVALID, "wf_record_face_samples: face mask has bits above 5");
    if (every == 0 || rows == 0) return fail(WF_ERR_INVALID, "wf_record_face_samples: every and rows must be >= 1");
    for (int f = 0; f < 6; ++f) {
        if (!(face_mask >> f & 1u)) continue;
        const FaceGeom fg = face_geom(c, f);
        const wf_face_region &r = regions[f];
        const int dir = fg.hi ? -1 : 1;
        const int deep = r.b0 + 3 * dir, ghost = r.b0 - dir;
        const bool ok = r.b0 >= 0 && r.b0 < fg.na && deep >= 0 && deep < fg.na && ghost >= 0 &&
                        ghost < fg.na && r.nfast > 0 && r.nslow > 0 && r.fast0 >= 0 &&
                        r.slow0 >= 0 && r.fast0 + r.nfast <= fg.nf && r.slow0 + r.nslow <= fg.ns;
        if (!ok) return fail(WF_ERR_INVALID, "face %d: region outside the domain", f);
    }
    WF_HIP(hipSetDevice(c->device));
    WF_HIP(hipStreamSynchronize(c->stream));
    drop_graph_exec(c);
    int rc = WF_OK;
    for (int f = 0; f < 6 && !rc; ++f) {
        if (!(face_mask >> f & 1u)) continue;
        clear_face(c, f);
        const FaceGeom fg = face_geom(c, f);
        const wf_face_region &r = regions[f];
        const uint64_t npos = (uint64_t)r.nfast * r.nslow;
        const uint64_t nsamp = rows * 2 * npos * kSampleStride;
        rc = dmalloc(&c->d_mem[f], 2 * npos * kMemStride * sizeof(float));
        if (!rc) rc = dmalloc(&c->d_samples[f], nsamp * sizeof(float));
        if (rc) break;
        if (hipMemset(c->d_mem[f], 0, 2 * npos * kMemStride * sizeof(float)) != hipSuccess ||
            hipMemset(c->d_samples[f], 0, nsamp * sizeof(float)) != hipSuccess) {
            rc = fail(WF_ERR_HIP, "wf_record_face_samples: clear failed");
            break;
        }
        FaceHook &h = c->hook[f];
        h.axis = fg.a;
        h.hi = fg.hi;
        h.b0 = r.b0;
        h.fast0 = r.fast0;
        h.nfast = r.nfast;
        h.slow0 = r.slow0;
        h.nslow = r.nslow;
        h.na = fg.na;
        h.nf = fg.nf;
        h.ns = fg.ns;
        h.mode = 1;
        h.mem = c->d_mem[f];
        h.samples = c->d_samples[f];
        h.emit_base = c->step;
        h.emit_every = every;
        h.emit_rows = rows;
        c->hook_on[f] = true;
    }
    update_ghost_reads(c);
    return rc;
}

int wf_replay_faces(wf_ctx *c, uint32_t face_mask, uint64_t rows, const float *const *samples)
{
    clear_error();
    if (!c || (face_mask && !samples)) return fail(WF_ERR_INVALID, "wf_replay_faces: null pointer");
    if (face_mask >= 64u) return fail(WF_ERR_INVALID, "wf_replay_faces: face mask has bits above 5");
    if (rows == 0) return fail(WF_ERR_INVALID, "wf_replay_faces: rows must be >= 1");
    for (int f = 0; f < 6; ++f) {
        if (!(face_mask >> f & 1u)) continue;
        if (!samples[f]) return fail(WF_ERR_INVALID, "wf_replay_faces: face %d has no samples", f);
        if (c->g.periodic >> (f / 2) & 1u) return fail(WF_ERR_INVALID, "face %d lies on a periodic axis", f);
        if (c->pml && c->g.T[f / 2] < 4)
            return fail(WF_ERR_INVALID, "face %d: in a domain with a CPML its axis needs a shell "
                        "of at least 4 cells", f);
    }
    WF_HIP(hipSetDevice(c->device));
    WF_HIP(hipStreamSynchronize(c->stream));
    drop_graph_exec(c);
    int rc = WF_OK;
    for (int f = 0; f < 6 && !rc; ++f) {
        if (!(face_mask >> f & 1u)) continue;
        clear_face(c, f);
        const FaceGeom fg = face_geom(c, f);
        const uint64_t npos = (uint64_t)fg.nf * fg.ns;
        const uint64_t nsamp = rows * 2 * npos * kSampleStride;
        rc = dmalloc(&c->d_ghost[f], 2 * npos * sizeof(float));
        if (!rc) rc = dmalloc(&c->d_samples[f], nsamp * sizeof(float));
        if (rc) break;
        if (hipMemset(c->d_ghost[f], 0, 2 * npos * sizeof(float)) != hipSuccess ||
            hipMemcpy(c->d_samples[f], samples[f], nsamp * sizeof(float), hipMemcpyHostToDevice) !=
                hipSuccess) {
            rc = fail(WF_ERR_HIP, "wf_replay_faces: upload failed");
            break;
        }
        FaceHook &h = c->hook[f];
        h.axis = fg.a;
        h.hi = fg.hi;
        h.b0 = fg.hi ? fg.na - 1 : 0;
        h.nfast = h.nf = fg.nf;
        h.nslow = h.ns = fg.ns;
        h.na = fg.na;
        h.mode = 2;
        h.ghost[0] = c->d_ghost[f];
        h.ghost[1] = c->d_ghost[f] + npos;
        h.samples = c->d_samples[f];
        h.emit_base = c->step;
        h.emit_every = 1;
        h.emit_rows = rows;
        for (int q = 0; q < 2; ++q) c->p.ghost[f][q] = c->alt.ghost[f][q] = h.ghost[q];
        c->hook_on[f] = true;
    }
    update_ghost_reads(c);
    return rc;
}

int wf_read_face_state(wf_ctx *c, uint32_t face, float *dst)
{
    clear_error();
    if (!c || !dst) return fail(WF_ERR_INVALID, "wf_read_face_state: null pointer");
    if (face >= 6 || !c->hook_on[face] || c->hook[face].mode == 1 || !c->d_mem[face])
        return fail(WF_ERR_STATE, "face %u has no predicted ghost plane", face);
    const FaceHook &h = c->hook[face];
    const size_t n = 2ull * h.nfast * h.nslow;
    std::vector<float> m(n * kMemStride);
    WF_HIP(hipSetDevice(c->device));
    WF_HIP(hipStreamSynchronize(c->stream));
    WF_HIP(hipMemcpy(m.data(), c->d_mem[face], m.size() * sizeof(float), hipMemcpyDeviceToHost));
    // g(n-1), g(n-2), P0(n-1), P0(n-2), then the four statistics
    for (size_t i = 0; i < n; ++i) {
        std::memcpy(dst + 8 * i, &m[i * kMemStride], 4 * sizeof(float));
        std::memcpy(dst + 8 * i + 4, &m[i * kMemStride + kMemStats], 4 * sizeof(float));
    }
    return WF_OK;
}

uint64_t wf_face_sample_count(const wf_ctx *c, uint32_t face)
{
    if (!c || face >= 6 || !c->hook_on[face] || c->hook[face].mode != 1) return 0;
    const FaceHook &h = c->hook[face];
    return h.emit_rows * 2ull * (uint64_t)h.nfast * h.nslow;
}

int wf_read_face_samples(wf_ctx *c, uint32_t face, float *dst)
{
    clear_error();
    if (!c || !dst) return fail(WF_ERR_INVALID, "wf_read_face_samples: null pointer");
    const uint64_t n = wf_face_sample_count(c, face);
    if (n == 0) return fail(WF_ERR_STATE, "face %u is not recording samples", face);
    WF_HIP(hipSetDevice(c->device));
    WF_HIP(hipStreamSynchronize(c->stream));
    WF_HIP(hipMemcpy(dst, c->d_samples[face], n * kSampleStride * sizeof(float), hipMemcpyDeviceToHost));
    return WF_OK;
}

// ===========================================================================
// Batches: many scenes advanced together, one launch per phase
// ===========================================================================
}  // extern "C"

struct wf_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;  // end of the last batched steps
    std::vector<wf_ctx *> scenes;
    BatchScene *d_scenes = nullptr;
    uint32_t *d_wg = nullptr;
    uint64_t wg_cap = 0;
    int total = 0;
    int zchunk = 8;
    int lpr_cap = 64;
    int grid = 0;  // 0: one workgroup per tile; else persistent workgroups
    int tiled_zchunk = -1, tiled_lpr = -1;  // tiling the wg table was built for
    std::vector<Box> boxes;
    // two steps per pass (every scene fused, no periodic axis; WF_BATCH_TWOSTEP=0
    // disables): each scene's regions as items, grouped by CPML variant, one
    // launch per variant; item tables for both state parities (a = the
    // context's current state, then swapped), as the scene descriptors
    bool pairs = false;
    std::vector<PassItem> plan;
    int var_first[8] = {}, var_total[8] = {};  // workgroup table offset, tiles per variant
    TwoStepItem *d_items = nullptr;            // [2][plan.size()]
    uint32_t *d_wgi = nullptr;                 // workgroup -> item
    uint64_t pass_bytes = 0;                   // algorithmic bytes of one pass
    int slice = 0;                             // WF_BATCH_SLICE: planes per z slice (0: modelled)
    int plan_mode = -1;  // WF_BATCH_PLAN: -1 by width, 0 regions, 1 full rows, 2 whole scenes
    int nstreams = 4;                          // WF_BATCH_STREAMS: variant launches spread over
    hipStream_t vs[4] = {};                    // the extra streams
    hipEvent_t fork = nullptr, join[4] = {};
};

namespace {

// The batch's items: every scene's, grouped by variant.
void batch_plan(wf_batch *b)
{
    std::vector<PassItem> items;
    std::vector<const Geo *> geo;
    for (size_t i = 0; i < b->scenes.size(); ++i) {
        pass_items(b->scenes[i]->g, b->plan_mode, (int)i, items);
        geo.push_back(&b->scenes[i]->g);
    }
    b->pass_bytes = group_items(items, geo, b->slice, b->scenes[0]->cus, b->plan, b->var_first,
                                b->var_total);
}

// Scene descriptors from the contexts' current state and tables; the
// workgroup -> scene table is rebuilt when the tiling changes.
int batch_upload(wf_batch *b)
{
    const size_t n = b->scenes.size();
    if (b->tiled_zchunk != b->zchunk || b->tiled_lpr != b->lpr_cap) {
        b->boxes.resize(n);
        uint64_t total = 0;
        for (size_t i = 0; i < n; ++i) {
            const Geo &g = b->scenes[i]->g;
            Box bx = split_box(0, g.nx, 0, g.ny, 0, g.nz, b->zchunk, b->lpr_cap);
            bx.first = (int)total;
            total += (uint64_t)bx.gx * bx.gy * bx.gz;
            b->boxes[i] = bx;
        }
        if (total > 0x7fffffffull) return fail(WF_ERR_INVALID, "batch: too many workgroups");
        std::vector<uint32_t> wg(total);
        for (size_t i = 0; i < n; ++i) {
            const Box &bx = b->boxes[i];
            const uint64_t cnt = (uint64_t)bx.gx * bx.gy * bx.gz;
            for (uint64_t k = 0; k < cnt; ++k) wg[bx.first + k] = (uint32_t)i;
        }
        if (total > b->wg_cap) {
            dfree(b->d_wg);
            b->wg_cap = 0;
            if (int rc = dmalloc(&b->d_wg, total * sizeof(uint32_t))) return rc;
            b->wg_cap = total;
        }
        if (total)
            WF_HIP(hipMemcpy(b->d_wg, wg.data(), total * sizeof(uint32_t), hipMemcpyHostToDevice));
        b->total = (int)total;
        b->tiled_zchunk = b->zchunk;
        b->tiled_lpr = b->lpr_cap;
    }
    // descriptors of both parities: [0] on the current state, [1] on the
    // other one (the current state after an odd number of two-step passes)
    std::vector<BatchScene> h(2 * n);
    for (size_t i = 0; i < 2 * n; ++i) {
        const wf_ctx *c = b->scenes[i % n];
        BatchScene &d = h[i];
        std::memset(&d, 0, sizeof d);
        d.g = c->g;
        d.p = i < n ? c->p : c->alt;
        d.slot = c->d_pslot;
        d.box = b->boxes[i % n];
        d.src = c->d_src;
        d.src_grp = c->d_src_grp;
        d.nsrc_grp = (int)c->nsrc_grp;
        d.srcval = c->d_srcval;
        d.src_base = c->src_base;
        d.nsrc = (int)c->nsrc;
        d.ncol = (int)c->nsrc_col;
        d.probe = c->d_probe;
        d.cos_tab = c->d_cos;
        d.sin_tab = c->d_sin;
        d.active = c->d_active;
        d.sums = c->d_sums;
        d.counts = c->d_counts;
        d.trace = c->d_trace;
        d.probe_base = c->probe_base;
        d.nprobe = (int)c->nprobe;
        d.nfreq = (int)c->nfreq;
        d.step = c->d_step;
    }
    WF_HIP(hipMemcpy(b->d_scenes, h.data(), 2 * n * sizeof(BatchScene), hipMemcpyHostToDevice));
    if (b->pairs) {
        // the regions' items of both parities: [0] reads the current state
        // and writes the other, [1] the reverse
        const size_t m = b->plan.size();
        std::vector<TwoStepItem> it(2 * m);
        for (size_t k = 0; k < 2 * m; ++k) {
            const PassItem &pi = b->plan[k % m];
            const wf_ctx *c = b->scenes[pi.scene];
            TwoStepItem &d = it[k];
            std::memset(&d, 0, sizeof d);
            d.g = c->g;
            d.a = k < m ? c->p : c->alt;
            d.b = k < m ? c->alt : c->p;
            d.box = pi.box;
            d.pts.src = PointRows{c->src_pts.bits, c->src_pts.key, c->src_pts.ent, c->src_pts.n};
            d.pts.prb = PointRows{c->prb_pts.bits, c->prb_pts.key, c->prb_pts.ent, c->prb_pts.n};
            d.pts.values = c->d_srcval;
            d.pts.ncol = (int)c->nsrc_col;
            d.pts.base = c->src_base;
            d.pts.step = c->d_step;
            d.pts.slot = c->d_pslot;
        }
        WF_HIP(hipMemcpy(b->d_items, it.data(), 2 * m * sizeof(TwoStepItem), hipMemcpyHostToDevice));
    }
    return WF_OK;
}

int batch_launch_steps(wf_batch *b, uint64_t nsteps)
{
    const unsigned n = (unsigned)b->scenes.size();
    uint64_t s = 0, passes = 0;
    if (b->pairs) {
        // two loop iterations per pass: the two-step kernels of every
        // variant (state A -> B, the first step's sources and probe values
        // in-kernel), then every scene's remaining point work on B
        const size_t m = b->plan.size();
        for (; nsteps - s >= 2; s += 2, ++passes) {
            const TwoStepItem *items = b->d_items + (passes & 1) * m;
            const int ns = b->nstreams;
            if (ns > 1) {
                WF_HIP(hipEventRecord(b->fork, b->stream));
                for (int k = 1; k < ns; ++k) WF_HIP(hipStreamWaitEvent(b->vs[k], b->fork, 0));
            }
            for (int v = 0, j = 0; v < 8; ++v) {
                if (!b->var_total[v]) continue;
                hipStream_t st = (j % ns) ? b->vs[j % ns] : b->stream;
                ++j;
                WF_HIP(launch_batch_twostep(items, b->d_wgi + b->var_first[v], b->var_total[v], v, st));
            }
            for (int k = 1; k < ns; ++k) {
                WF_HIP(hipEventRecord(b->join[k], b->vs[k]));
                WF_HIP(hipStreamWaitEvent(b->stream, b->join[k], 0));
            }
            hipLaunchKernelGGL(k_pair_points_batch, dim3(n), dim3(64), 0, b->stream,
                               b->d_scenes + ((passes + 1) & 1) * n);
            WF_HIP(hipGetLastError());
        }
    }
    const BatchScene *sc = b->d_scenes + (passes & 1) * n;
    for (; s < nsteps; ++s) {
        WF_HIP(launch_batch_h(sc, b->d_wg, b->total, b->zchunk, b->grid, b->stream));
        WF_HIP(launch_batch_e(sc, b->d_wg, b->total, b->zchunk, b->grid, b->stream));
        hipLaunchKernelGGL(k_sources_batch, dim3(n), dim3(64), 0, b->stream, sc);
        hipLaunchKernelGGL(k_probes_batch, dim3(n), dim3(64), 0, b->stream, sc);
        WF_HIP(hipGetLastError());
    }
    // later work queued on a scene's own stream runs after these steps
    WF_HIP(hipEventRecord(b->done, b->stream));
    for (wf_ctx *c : b->scenes) {
        WF_HIP(hipStreamWaitEvent(c->stream, b->done, 0));
        if (passes & 1) std::swap(c->p, c->alt);
        c->step += nsteps;
    }
    return WF_OK;
}

// Checks the tables, waits for each scene's own queued work and uploads the
// descriptors.
int batch_prepare(wf_batch *b, uint64_t nsteps)
{
    for (size_t i = 0; i < b->scenes.size(); ++i) {
        const wf_ctx *c = b->scenes[i];
        for (int f = 0; f < 6; ++f)
            if (c->hook_on[f])
                return fail(WF_ERR_STATE, "batch scene %zu has learned-boundary faces (step it alone)", i);
        if (c->nsrc && c->step + nsteps > c->src_base + c->src_steps)
            return fail(WF_ERR_STATE, "batch scene %zu: source table covers %llu steps", i,
                        (unsigned long long)c->src_steps);
        if (c->nprobe && c->step + nsteps > c->probe_base + c->probe_steps)
            return fail(WF_ERR_STATE, "batch scene %zu: probe table covers %llu steps", i,
                        (unsigned long long)c->probe_steps);
        if (b->pairs && ((c->nsrc && !c->src_pts.bits) || (c->nprobe && (!c->prb_pts.bits || !c->d_pslot))))
            return fail(WF_ERR_STATE, "batch scene %zu: no two-step point tables", i);
    }
    WF_HIP(hipSetDevice(b->device));
    for (wf_ctx *c : b->scenes) WF_HIP(hipStreamSynchronize(c->stream));
    // the descriptor and workgroup tables are rewritten by synchronous copies,
    // which do not wait for the (non-blocking) batch stream: earlier batched
    // steps must be done reading them
    WF_HIP(hipStreamSynchronize(b->stream));
    return batch_upload(b);
}

}  // namespace

extern "C" {

int wf_batch_create(uint32_t n, wf_ctx *const *scenes, wf_batch **out)
{
    clear_error();
    if (!out || (n && !scenes)) return fail(WF_ERR_INVALID, "wf_batch_create: null pointer");
    *out = nullptr;
    if (n == 0) return fail(WF_ERR_INVALID, "wf_batch_create: empty batch");
    for (uint32_t i = 0; i < n; ++i) {
        if (!scenes[i]) return fail(WF_ERR_INVALID, "wf_batch_create: scene %u is null", i);
        if (scenes[i]->device != scenes[0]->device)
            return fail(WF_ERR_INVALID, "wf_batch_create: scenes on different devices");
        for (uint32_t j = 0; j < i; ++j)
            if (scenes[j] == scenes[i])
                return fail(WF_ERR_INVALID, "wf_batch_create: scene %u listed twice", i);
    }
    wf_batch *b = new wf_batch;
    b->device = scenes[0]->device;
    b->scenes.assign(scenes, scenes + n);
    if (const char *z = std::getenv("WF_BATCH_ZCHUNK")) b->zchunk = std::max(1, std::atoi(z));
    if (const char *z = std::getenv("WF_BATCH_SLICE")) b->slice = std::max(0, std::atoi(z));
    if (const char *z = std::getenv("WF_BATCH_PLAN")) b->plan_mode = std::min(2, std::max(-1, std::atoi(z)));
    if (const char *z = std::getenv("WF_BATCH_STREAMS")) b->nstreams = std::min(4, std::max(1, std::atoi(z)));
    const char *tw = std::getenv("WF_BATCH_TWOSTEP");
    b->pairs = !(tw && tw[0] == '0');
    for (const wf_ctx *c : b->scenes)
        if (!pair_tables(c) || !c->d_fields2) b->pairs = false;
    int rc = WF_OK;
    if (hipSetDevice(b->device) != hipSuccess ||
        hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess)
        rc = fail(WF_ERR_HIP, "wf_batch_create: stream creation failed");
    if (!rc) rc = dmalloc(&b->d_scenes, 2ull * n * sizeof(BatchScene));
    if (!rc && b->nstreams > 1) {
        if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
            rc = fail(WF_ERR_HIP, "wf_batch_create: event creation failed");
        for (int k = 1; k < b->nstreams && !rc; ++k)
            if (hipStreamCreateWithFlags(&b->vs[k], hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming) != hipSuccess)
                rc = fail(WF_ERR_HIP, "wf_batch_create: stream creation failed");
    }
    if (!rc && b->pairs) {
        batch_plan(b);
        uint64_t tiles = 0;
        for (int v = 0; v < 8; ++v) tiles += (uint64_t)b->var_total[v];
        if (tiles > 0x7fffffffull) rc = fail(WF_ERR_INVALID, "batch: too many workgroups");
        std::vector<uint32_t> wg(tiles);
        for (size_t k = 0; k < b->plan.size(); ++k) {
            const PassItem &it = b->plan[k];
            const uint64_t cnt = (uint64_t)it.box.gx * it.box.gy * it.box.gz;
            for (uint64_t j = 0; j < cnt; ++j) wg[b->var_first[it.axes] + it.box.first + j] = (uint32_t)k;
        }
        if (!rc) rc = dmalloc(&b->d_items, 2 * b->plan.size() * sizeof(TwoStepItem));
        if (!rc) rc = dmalloc(&b->d_wgi, std::max<uint64_t>(tiles, 1) * sizeof(uint32_t));
        if (!rc && tiles &&
            hipMemcpy(b->d_wgi, wg.data(), tiles * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(WF_ERR_HIP, "wf_batch_create: table upload failed");
    }
    if (rc) {
        wf_batch_destroy(b);
        return rc;
    }
    *out = b;
    return WF_OK;
}

void wf_batch_destroy(wf_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    dfree(b->d_scenes);
    dfree(b->d_wg);
    dfree(b->d_items);
    dfree(b->d_wgi);
    for (int k = 1; k < 4; ++k) {
        if (b->vs[k]) {
            (void)hipStreamSynchronize(b->vs[k]);
            (void)hipStreamDestroy(b->vs[k]);
        }
        if (b->join[k]) (void)hipEventDestroy(b->join[k]);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int wf_batch_step(wf_batch *b, uint64_t nsteps)
{
    clear_error();
    if (!b) return fail(WF_ERR_INVALID, "wf_batch_step: null batch");
    if (int rc = batch_prepare(b, nsteps)) return rc;
    return batch_launch_steps(b, nsteps);
}

int wf_batch_sync(wf_batch *b)
{
    if (!b) return fail(WF_ERR_INVALID, "wf_batch_sync: null batch");
    WF_HIP(hipSetDevice(b->device));
    WF_HIP(hipStreamSynchronize(b->stream));
    return WF_OK;
}

int wf_batch_time_steps(wf_batch *b, uint64_t nsteps, double *total_ms)
{
    clear_error();
    if (!b || !total_ms) return fail(WF_ERR_INVALID, "wf_batch_time_steps: null pointer");
    if (int rc = batch_prepare(b, nsteps)) return rc;
    hipEvent_t e0, e1;
    WF_HIP(hipEventCreate(&e0));
    WF_HIP(hipEventCreate(&e1));
    WF_HIP(hipEventRecord(e0, b->stream));
    int rc = batch_launch_steps(b, nsteps);
    WF_HIP(hipEventRecord(e1, b->stream));
    WF_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    WF_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *total_ms = ms;
    return rc;
}

int wf_batch_tune(wf_batch *b, uint32_t steps_per_candidate, int *zchunk, int *lanes_per_row,
                  int *grid, uint64_t *steps_taken)
{
    clear_error();
    if (!b) return fail(WF_ERR_INVALID, "wf_batch_tune: null batch");
    if (steps_per_candidate < 2)
        return fail(WF_ERR_INVALID, "wf_batch_tune: need at least 2 steps per candidate");
    // (z-chunk, lanes-per-row cap, grid): short z-marches fill the chip with
    // small scenes; a 16-lane cap avoids idle lanes on rows of 33-48 quads;
    // a persistent grid (8 or 32 workgroups per CU) trades launch width for
    // fewer, longer-lived workgroups
    // lanes-per-row cap 0: rows fitted to the scene's width (split_box)
    static const int cand[][3] = {{4, 64, 0},    {8, 64, 0},    {16, 64, 0},
                                  {32, 64, 0},   {8, 16, 0},    {16, 16, 0},
                                  {8, 64, 2048}, {16, 64, 2048}, {8, 64, 8192},
                                  {2, 0, 0},     {4, 0, 0},     {8, 0, 0},
                                  {4, 0, 2048}};
    const int ncand = (int)(sizeof cand / sizeof cand[0]);
    if (b->pairs) {
        // two-step passes: the regions follow each scene's geometry and the
        // tiling above applies only to a trailing single step; no steps taken
        if (zchunk) *zchunk = b->zchunk;
        if (lanes_per_row) *lanes_per_row = b->lpr_cap;
        if (grid) *grid = b->grid;
        if (steps_taken) *steps_taken = 0;
        return WF_OK;
    }
    if (int rc = batch_prepare(b, (uint64_t)ncand * steps_per_candidate)) return rc;
    hipEvent_t e0, e1;
    WF_HIP(hipEventCreate(&e0));
    WF_HIP(hipEventCreate(&e1));
    double best = 0.0;
    int bz = b->zchunk, bl = b->lpr_cap, bg = b->grid, rc = WF_OK;
    uint64_t taken = 0;
    for (int k = 0; k < ncand && !rc; ++k) {
        // every trial step is a real step: results do not depend on tiling
        b->zchunk = cand[k][0];
        b->lpr_cap = cand[k][1];
        b->grid = cand[k][2];
        rc = batch_upload(b);
        if (!rc) rc = batch_launch_steps(b, 1);
        if (!rc && hipEventRecord(e0, b->stream) != hipSuccess) rc = fail(WF_ERR_HIP, "event");
        if (!rc) rc = batch_launch_steps(b, steps_per_candidate - 1);
        if (!rc && hipEventRecord(e1, b->stream) != hipSuccess) rc = fail(WF_ERR_HIP, "event");
        float ms = 0.f;
        if (!rc && (hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = fail(WF_ERR_HIP, "wf_batch_tune: event timing failed");
        taken += steps_per_candidate;
        if (!rc && (k == 0 || ms < best)) {
            best = ms;
            bz = cand[k][0];
            bl = cand[k][1];
            bg = cand[k][2];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->zchunk = bz;
    b->lpr_cap = bl;
    b->grid = bg;
    if (zchunk) *zchunk = bz;
    if (lanes_per_row) *lanes_per_row = bl;
    if (grid) *grid = bg;
    if (steps_taken) *steps_taken = taken;
    return rc;
}

int wf_batch_set_tiling(wf_batch *b, int zchunk, int lanes_per_row, int grid)
{
    if (!b) return fail(WF_ERR_INVALID, "wf_batch_set_tiling: null batch");
    if (zchunk < 1 || lanes_per_row < 0 || lanes_per_row > 64 ||
        (lanes_per_row & (lanes_per_row - 1)) || grid < 0)
        return fail(WF_ERR_INVALID,
                    "wf_batch_set_tiling: zchunk >= 1, lanes per row a power of two <= 64, grid >= 0");
    b->zchunk = zchunk;
    b->lpr_cap = lanes_per_row;
    b->grid = grid;
    return WF_OK;
}

int wf_batch_two_step(const wf_batch *b, int *two_step, uint64_t *pass_bytes)
{
    if (!b || !two_step || !pass_bytes) return fail(WF_ERR_INVALID, "wf_batch_two_step: null pointer");
    *two_step = b->pairs ? 1 : 0;
    *pass_bytes = b->pairs ? b->pass_bytes : 0;
    return WF_OK;
}

uint64_t wf_batch_cells(const wf_batch *b)
{
    if (!b) return 0;
    uint64_t n = 0;
    for (const wf_ctx *c : b->scenes) n += (uint64_t)c->g.nx * c->g.ny * c->g.nz;
    return n;
}

}  // extern "C"

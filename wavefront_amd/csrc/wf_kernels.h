// Device-side data structures and the per-voxel Yee/CPML update shared by
// every kernel variant.  Included by HIP translation units only.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace wf {

// Grid geometry and boundary description, passed by value to kernels.
struct Geo {
    int nx, ny, nz;
    int nb[3];        // blocks per axis (n/8)
    int T[3];         // graded CPML depth per axis, voxels (0 = none)
    int sb[3];        // PML shell depth per axis, blocks (0 = none)
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
    const float4 *coef;  // 256 x (ca, cb, da, db), dequantized Q16.16
    const float *prof[3][3];  // [axis][b, a, inv_kappa], length n_axis
    float *psi[3][4];
};

struct KArgs {
    Geo g;
    Ptrs p;
};

// Source / probe tables.
struct DevSource {
    uint64_t off;  // voxel linear index
    int comp;      // 0..2
    int hard;
};

struct DevProbe {
    uint64_t off;
    int comp;  // 0..5
    int nfreq;
    int foff;  // offset into the flattened frequency axis
    int pad;
};

// One-tap recursive convolution, src/fdtd.rs:136-163:
//   psi = b*psi + a*raw;  d = raw*inv_kappa + psi
__device__ __forceinline__ float cpml(float raw, float b, float a, float ik, float &psi)
{
    const float pn = b * psi + a * raw;
    psi = pn;
    return raw * ik + pn;
}

}  // namespace wf

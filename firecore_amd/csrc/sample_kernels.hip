// Sampling kernels behind the reference's diagnostic entry points
// (libMMFF.h:439-497 sample_*, MMFFmulti_lib.cpp:168-250 sampleSurf*): they
// evaluate the device's own FP32 force-field terms on host-supplied sample
// points so that the terms can be compared with the CPU definitions.
#include "kernels.h"

namespace fcd {

// kind 0: evalPiAling, 1: evalAngleCos, 2: evalAngleCosHalf; h1=(1,0,0),
// h2=(cos a, sin a, 0); outputs E and f1.y (libMMFF.h:439-473).
__global__ void k_sample_angles(int kind, float k, float ang0, float r1, float r2, int n, const float* ang, float* E,
                                float* F) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const V3 h1 = mk(1.f, 0.f, 0.f);
    const V3 h2 = mk(cosf(ang[i]), sinf(ang[i]), 0.f);
    V3 f1, f2;
    float e = 0.f;
    if (kind == 0) {
        e = pi_align(h1, h2, k, f1, f2);
    } else if (kind == 1) {
        e = angle_cos(h1, h2, 1.f / r1, 1.f / r2, k, cosf(ang0), f1, f2);
    } else {
        e = angle_cos_half(h1, h2, 1.f / r1, 1.f / r2, cosf(ang0 * 0.5f), sinf(ang0 * 0.5f), k, f1, f2);
    }
    E[i] = e;
    F[i] = f1.y;
}

// GridFF sample of a probe with coefficients plq at positions pos[i]; out = (F, E).
__global__ void k_sample_grid(GridDev g, int n, const float4* pos, float4 plq, float4* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V3 f = zero3();
    const float e = grid_sample(g, xyz(pos[i]), mk(plq.x, plq.y, plq.z), f);
    out[i] = make_float4(f.x, f.y, f.z, e);
}

// Directed pair term getLJQH at distance r along x; out = (F_x on j side, E).
__global__ void k_sample_pair(int n, const float* r, float4 REQij, float2* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float fr, e;
    const V3 d = mk(r[i], 0.f, 0.f);
    lj_pair(d, REQij.x, REQij.y, COULOMB_K * REQij.z, REQij.w, fr, e);
    out[i] = make_float2(d.x * fr, e);  // force on the atom at +r along +x
}

}  // namespace fcd

extern "C" hipError_t fc_sample_angles(int kind, float k, float ang0, float r1, float r2, int n, const float* ang,
                                       float* E, float* F) {
    hipLaunchKernelGGL(fcd::k_sample_angles, dim3((n + 63) / 64), dim3(64), 0, 0, kind, k, ang0, r1, r2, n, ang, E, F);
    return hipGetLastError();
}
extern "C" hipError_t fc_sample_grid(const fcd::GridDev* g, int n, const float4* pos, float4 plq, float4* out) {
    hipLaunchKernelGGL(fcd::k_sample_grid, dim3((n + 63) / 64), dim3(64), 0, 0, *g, n, pos, plq, out);
    return hipGetLastError();
}
extern "C" hipError_t fc_sample_pair(int n, const float* r, float4 REQij, float2* out) {
    hipLaunchKernelGGL(fcd::k_sample_pair, dim3((n + 63) / 64), dim3(64), 0, 0, n, r, REQij, out);
    return hipGetLastError();
}

// GridFF node potentials on the device (FP64): one thread per grid node sums
// the Pauli and London channels over the Morse images and the bare Coulomb
// channel over the Coulomb images of every substrate atom -- the same sums,
// in the same order, as the host definition fc::gridff_point (node potentials
// of the reference's GridFF::evalBsplineRef, GridFF.h:848-890, with
// evalGridFFPoint_Mors :582-611 and evalGridFFPoint_Coul :614-635).  The
// B-spline fit of the node values stays on the host (three passes of
// tridiagonal solves, linear in the node count).
#include <hip/hip_runtime.h>

#include "fc_host.h"

namespace fcd {

struct GridGenDev {
    int nx, ny, nz, na;
    double p0x, p0y, p0z, sx, sy, sz;
    double ax, ay, bx, by;  // lateral lattice vectors
    double alpha;
    int nmx, nmy, ncx, ncy;  // Morse and Coulomb image ranges
    double rdamp2;           // Coulomb damping Rdamp^2 (0 = bare, r^2 bounded below)
};

__global__ void __launch_bounds__(256) k_gridff_nodes(GridGenDev G, const double* __restrict__ apos,
                                                      const double* __restrict__ REQs, double* __restrict__ V) {
    const long node = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long np = (long)G.nx * G.ny * G.nz;
    if (node >= np) return;
    const int iz = (int)(node % G.nz);
    const int iy = (int)((node / G.nz) % G.ny);
    const int ix = (int)(node / ((long)G.nz * G.ny));
    const double px = G.p0x + ix * G.sx, py = G.p0y + iy * G.sy, pz = G.p0z + iz * G.sz;
    constexpr double COULOMB_K = 14.3996448915;  // eV*A/e^2 (cpp/common/constants.h:9)
    double P = 0.0, L = 0.0, Q = 0.0;
    for (int a = 0; a < G.na; a++) {
        const double Ax = apos[a * 3], Ay = apos[a * 3 + 1], Az = apos[a * 3 + 2];
        const double R0 = REQs[a * 4], E = REQs[a * 4 + 1], q = REQs[a * 4 + 2];
        const double dz = pz - Az;
        for (int jy = -G.nmy; jy <= G.nmy; jy++)
            for (int jx = -G.nmx; jx <= G.nmx; jx++) {
                const double dx = px - Ax + jx * G.ax + jy * G.bx;
                const double dy = py - Ay + jx * G.ay + jy * G.by;
                const double r = sqrt(dx * dx + dy * dy + dz * dz + 1e-32);
                const double e = exp(-G.alpha * (r - R0));
                const double eM = e * E;
                P += eM * e;
                L += eM * -2.0;
            }
        const double kq = COULOMB_K * q;
        for (int jy = -G.ncy; jy <= G.ncy; jy++)
            for (int jx = -G.ncx; jx <= G.ncx; jx++) {
                const double dx = px - Ax + jx * G.ax + jy * G.bx;
                const double dy = py - Ay + jx * G.ay + jy * G.by;
                const double d2 = dx * dx + dy * dy + dz * dz;
                const double r2 = (G.rdamp2 > 0.0) ? d2 + G.rdamp2 : fmax(d2, fc::GRIDFF_R2MIN);
                Q += kq / sqrt(r2);
            }
    }
    V[node * 3] = P;
    V[node * 3 + 1] = L;
    V[node * 3 + 2] = Q;
}

}  // namespace fcd

// Node potentials of grid g over surface s, computed on the current device
// into V (host, nx*ny*nz*3, z fastest).  One-time set-up work: device buffers
// are allocated and released here.
extern "C" hipError_t fc_gridff_nodes_device(const fc::Surface* s, const fc::GridFFHost* g, double* V) {
    int nmx, nmy, ncx, ncy;
    fc::gridff_images(*s, nmx, nmy, ncx, ncy);
    fcd::GridGenDev G;
    G.nx = g->nx;
    G.ny = g->ny;
    G.nz = g->nz;
    G.na = s->natoms();
    G.p0x = g->pos0[0];
    G.p0y = g->pos0[1];
    G.p0z = g->pos0[2];
    G.sx = g->step[0];
    G.sy = g->step[1];
    G.sz = g->step[2];
    G.ax = s->lvec[0];
    G.ay = s->lvec[1];
    G.bx = s->lvec[3];
    G.by = s->lvec[4];
    G.alpha = g->alphaMorse;
    G.nmx = nmx;
    G.nmy = nmy;
    G.ncx = ncx;
    G.ncy = ncy;
    G.rdamp2 = s->Rdamp > 0.0 ? s->Rdamp * s->Rdamp : 0.0;
    const size_t np = (size_t)G.nx * G.ny * G.nz;
    if (np == 0 || G.na == 0) return hipErrorInvalidValue;
    double *d_apos = nullptr, *d_req = nullptr, *d_V = nullptr;
    hipError_t e = hipMalloc((void**)&d_apos, sizeof(double) * 3 * G.na);
    if (e == hipSuccess) e = hipMalloc((void**)&d_req, sizeof(double) * 4 * G.na);
    if (e == hipSuccess) e = hipMalloc((void**)&d_V, sizeof(double) * 3 * np);
    if (e == hipSuccess) e = hipMemcpy(d_apos, s->apos.data(), sizeof(double) * 3 * G.na, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_req, s->REQs.data(), sizeof(double) * 4 * G.na, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const unsigned nb = (unsigned)((np + 255) / 256);
        hipLaunchKernelGGL(fcd::k_gridff_nodes, dim3(nb), dim3(256), 0, 0, G, d_apos, d_req, d_V);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(V, d_V, sizeof(double) * 3 * np, hipMemcpyDeviceToHost);
    if (d_V) hipFree(d_V);
    if (d_req) hipFree(d_req);
    if (d_apos) hipFree(d_apos);
    return e;
}

// Split-phase kernels: one launch per phase of a FIRE step, over all live
// systems of the batch, in the reference's packed buffer layout
// (OCL_MM.h:197-250 / MolWorld_sp3_multi.h:514-592).  They mirror the
// reference's kernel decomposition (getNonBond_GridFF_Bspline, getMMFFf4,
// updateAtomsMMFFf4 in relax_multi.cl) but carry the CPU path's physics and a
// full per-system FIRE controller on the device.  They are the parity
// reference for the persistent kernel and the HBM-bound form of the path.
#include "kernels.h"
#include "precond.h"

namespace fcd {

// ---------------------------------------------------------------------------
// Non-bonded force on atom i of a periodic molecule, NBFF::
// evalLJQs_ng4_PBC_atom_omp (NBFF.h:575-620): every other atom j at every image
// shift, a bonded neighbour skipped only at the image of its bond (neighCell).
// The 1-3 pairs are then subtracted once per angle path at the image reached
// through the two bonds, as MMFFsp3_loc::eval_atom does in its angle loop
// (MMFFsp3_loc.h:525-536): the correction list of a periodic system holds one
// entry per path and end atom, dp = pj - pi + S[neighCell[c][kj]] -
// S[neighCell[c][ki]] (which may be an image outside the summed range, or the
// image of a 1-2 bond of the same two atoms in a short cell).  Adds the force
// to f and the unhalved energy to e.
__device__ void nb_pbc_atom(const BatchDev& B, const SysInfo& I, int i, V3 pi, float4 ri, V3& f, double& e) {
    const int N = I.natoms;
    const float kQi = COULOMB_K * ri.z;
    const float4* sh = B.pbcshifts + I.sh0;
    const int4 ng = B.neighs[I.a0 + i];
    const int4 nc = B.neighCell[I.a0 + i];
    for (int j = 0; j < N; j++) {
        if (j == i) continue;
        const V3 d0 = xyz(B.atoms[I.v0 + j]) - pi;
        const float4 rj = B.REQs[I.a0 + j];
        for (int s = 0; s < I.npbc; s++) {
            if ((j == ng.x && s == nc.x) || (j == ng.y && s == nc.y) || (j == ng.z && s == nc.z) ||
                (j == ng.w && s == nc.w))
                continue;
            const V3 d = d0 + xyz(sh[s]);
            float fr, ep;
            lj_pair<true, true>(d, ri.x + rj.x, ri.y * rj.y, kQi * rj.z, ri.w * rj.w, fr, ep);
            f -= d * fr;
            e += (double)ep;
        }
    }
    for (int c = 0; c < I.ncorr; c++) {
        const PairCorrDev pc = B.corr[I.corr0 + c];
        if (pc.i != i || pc.path < 0) continue;
        const int node = pc.path >> 4, ki = (pc.path >> 2) & 3, kj = pc.path & 3;
        const int* cell = reinterpret_cast<const int*>(B.neighCell + I.a0 + node);
        const int cj = cell[kj], ci = cell[ki];  // -1 only for non-finite positions at packing
        V3 d = xyz(B.atoms[I.v0 + pc.j]) - pi;
        if (cj >= 0) d += xyz(sh[cj]);
        if (ci >= 0) d -= xyz(sh[ci]);
        const float4 rj = B.REQs[I.a0 + pc.j];
        float fr, ep;
        lj_pair<true, true>(d, ri.x + rj.x, ri.y * rj.y, kQi * rj.z, ri.w * rj.w, fr, ep);
        f -= d * (fr * pc.w);
        e += (double)ep * pc.w;
    }
}

// ---------------------------------------------------------------------------
// Explicit substrate atoms, GridFF::evalMorsePBC (GridFF.h:1229-1261) with the
// pair term addAtomicForceMorseQ (Forces.h:540-561): for every substrate atom j
// (REQs mixed by combineREQ: R0 sum, E0 and charge products) at every image
// shift, dp = pi - pj - shift, r = |dp|, e = exp(K (r - R0)) with K =
// -alphaMorse; E = E0 e (e - 2) + k q q / r and the force on i is
// -dp (2 K E0 e (e - 1) / r - k q q / r^3).  Adds the force to f, returns E.
__device__ float surf_atoms_force(const RunParams& P, V3 pi, float4 ri, V3& f) {
    float E = 0.f;
    const float K = P.surf_K;
    for (int j = 0; j < P.nsurf; j++) {
        const float4 a = P.surf_pos[j];
        const float4 q = P.surf_REQ[j];
        const float r0 = q.x + ri.x, e0 = q.y * ri.y, kqq = COULOMB_K * q.z * ri.z;
        const V3 dp0 = mk(pi.x - a.x, pi.y - a.y, pi.z - a.z);
        for (int s = 0; s < P.nsurf_pbc; s++) {
            const V3 dp = dp0 + xyz(P.surf_shifts[s]);
            const float r2 = dot(dp, dp);
            const float ir = __builtin_amdgcn_rsqf(r2);
            const float r = r2 * ir;
            const float e = expf(K * (r - r0));
            const float Ae = e0 * e;
            const float eel = kqq * ir;
            E += Ae * (e - 2.f) + eel;
            const float fr = 2.f * K * Ae * (e - 1.f) * ir - eel * ir * ir;
            f -= dp * fr;
        }
    }
    return E;
}

// ---------------------------------------------------------------------------
// Non-bonded (+GridFF) force per atom.  grid = (ceil(maxN/64), nsys), 64 threads.
// aforce[v0+i] = (f, 0.5*E_nb + E_grid)   (assignment, like the reference NB kernel)
__global__ void __launch_bounds__(64) k_nonbond_grid(BatchDev B, RunParams P) {
    __shared__ float4 s_p[64];
    __shared__ float4 s_r[64];
    const int s = blockIdx.y;
    const SysInfo I = B.info[s];
    if (B.status[s] != ST_RUNNING) return;
    const int N = I.natoms;
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (blockIdx.x * 64 >= N) return;
    const bool act = i < N;
    const int W = (N + 31) >> 5;
    V3 pi = zero3();
    float4 ri = make_float4(0, 0, 0, 0);
    if (act) {
        pi = xyz(B.atoms[I.v0 + i]);
        ri = B.REQs[I.a0 + i];
    }
    const float kQi = COULOMB_K * ri.z;
    V3 f = zero3();
    double e = 0.0;  // FP64 accumulation of the pair energies
    if (P.bNonBonded && I.npbc > 0) {  // periodic molecule (uniform over the block)
        if (act) nb_pbc_atom(B, I, i, pi, ri, f, e);
        e *= 0.5;
    } else if (P.bNonBonded) {
        for (int j0 = 0; j0 < N; j0 += 64) {
            const int jl = j0 + threadIdx.x;
            __syncthreads();
            if (jl < N) {
                s_p[threadIdx.x] = B.atoms[I.v0 + jl];
                s_r[threadIdx.x] = B.REQs[I.a0 + jl];
            }
            __syncthreads();
            const int jn = min(64, N - j0);
            if (act) {
                unsigned m0 = B.excl[I.mask0 + i * W + (j0 >> 5)];
                unsigned m1 = ((j0 >> 5) + 1 < W) ? B.excl[I.mask0 + i * W + (j0 >> 5) + 1] : 0u;
                for (int jj = 0; jj < jn; jj++) {
                    const float4 pj = s_p[jj];
                    const float4 rj = s_r[jj];
                    const V3 d = mk(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z);
                    float fr, ep;
                    lj_pair<true, true>(d, ri.x + rj.x, ri.y * rj.y, kQi * rj.z, ri.w * rj.w, fr, ep);
                    const unsigned m = jj < 32 ? m0 : m1;
                    const bool ex = (m >> (jj & 31)) & 1u;
                    fr = ex ? 0.f : fr;
                    ep = ex ? 0.f : ep;
                    f.x -= d.x * fr;
                    f.y -= d.y * fr;
                    f.z -= d.z * fr;
                    e += (double)ep;
                }
            }
        }
        if (act) {
            for (int c = 0; c < I.ncorr; c++) {
                const PairCorrDev pc = B.corr[I.corr0 + c];
                if (pc.i != i) continue;
                const V3 pj = xyz(B.atoms[I.v0 + pc.j]);
                const float4 rj = B.REQs[I.a0 + pc.j];
                const V3 d = pj - pi;
                float fr, ep;
                lj_pair<true, true>(d, ri.x + rj.x, ri.y * rj.y, kQi * rj.z, ri.w * rj.w, fr, ep);
                f -= d * (fr * pc.w);
                e += (double)ep * pc.w;
            }
        }
        e *= 0.5;
    }
    if (!act) return;
    if (P.bGridFF) {
        const float4 plq = B.PLQs[I.a0 + i];
        e += (double)grid_sample(P.grid, pi, mk(plq.x, plq.y, plq.z), f);
    }
    if (P.nsurf > 0) e += (double)surf_atoms_force(P, pi, ri, f);
    B.aforce[I.v0 + i] = make_float4(f.x, f.y, f.z, 0.f);
    B.enb[I.a0 + i] = e;
}

// ---------------------------------------------------------------------------
// Bonded terms per node.  One thread per node of the batch; node_sys maps the
// global node index to its system.  Writes recoils into fneigh (sigma block
// then pi block of each system, reference layout), adds the node's own force
// to aforce[node] and assigns aforce[pi] (as getMMFFf4 does).
__global__ void __launch_bounds__(256) k_bonded(BatchDev B, RunParams P, const int* node_sys, int ntot_nodes) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ntot_nodes) return;
    const int s = node_sys[g];
    if (B.status[s] != ST_RUNNING) return;
    const SysInfo I = B.info[s];
    const int k = g - I.n0;
    const int N = I.natoms, M = I.nnode;
    const float4* apos = B.atoms + I.v0;
    auto pos = [&](int v) { return xyz(apos[v]); };
    // periodic molecules: neighbour k sits at its image pbc_shifts[neighCell[k]]
    const int4 nc = I.npbc > 0 ? B.neighCell[I.a0 + k] : make_int4(-1, -1, -1, -1);
    const float4* shifts = B.pbcshifts + I.sh0;
    auto nshift = [&](int q) {
        const int c = q == 0 ? nc.x : (q == 1 ? nc.y : (q == 2 ? nc.z : nc.w));
        return c >= 0 ? xyz(shifts[c]) : zero3();
    };
    V3 fbs[4], fps[4], fa, fpi;
    float E = eval_node(k, N, M, pos, nshift, B.neighs[I.a0 + k], B.apars[g], B.bLs[g], B.bKs[g], B.Ksp[g], B.Kpp[g],
                        P.sw, fbs, fps, fa, fpi);
    float4* fng = B.fneigh + (size_t)I.n0 * 8;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        fng[k * 4 + q] = make_float4(fbs[q].x, fbs[q].y, fbs[q].z, 0.f);
        fng[4 * M + k * 4 + q] = make_float4(fps[q].x, fps[q].y, fps[q].z, 0.f);
    }
    float4 a = B.aforce[I.v0 + k];
    a.x += fa.x;
    a.y += fa.y;
    a.z += fa.z;
    B.aforce[I.v0 + k] = a;
    B.aforce[I.v0 + N + k] = make_float4(fpi.x, fpi.y, fpi.z, 0.f);
    B.ebond[g] = E;
}

__device__ __forceinline__ float blk_sum(float v, float* sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    float r = 0.f;
    for (int w = 0; w < nw; w++) r += sh[w];
    return r;
}
__device__ __forceinline__ float blk_max(float v, float* sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    float r = sh[0];
    for (int w = 1; w < nw; w++) r = fmaxf(r, sh[w]);
    return r;
}
__device__ __forceinline__ double blk_sum_d(double v, double* sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0.0;
    for (int w = 0; w < nw; w++) r += sh[w];
    return r;
}

// ---------------------------------------------------------------------------
// Assemble + per-system reduction + FIRE + move.  One block (256) per system.
// do_move=0 only assembles (force/energy evaluation for parity checks).
// Move of the exploring form: the per-system phase machine of
// MolWorld_sp3_multi::updateMultiExploring (MolWorld_sp3_multi.h:770-833) run
// on the device, with the sums and energy of this step from k_update.
//   relaxing:  FIRE as in k_update; on convergence the minimum is recorded
//              (energy at this geometry, positions when minpos is set) and the
//              exploring phase starts; after xn_relax steps (> 0) it starts
//              without a record;
//   exploring: xn_explore steps of Langevin dynamics at dt_max with no FIRE
//              mixing (MDpars = (dt_max, 1, 1, 0), evalVF_new :755-760; drive
//              relax_multi.cl:1160-1178), then FIRE restarts from its initial
//              state (dt_max, damp_max).
// The run ends after nstep_max steps (status ST_MAXSTEP).
__device__ void explore_move(const BatchDev& B, const RunParams& P, int s, const SysInfo& I, int step, float vf,
                             float vv, float ff, float fm2, double e) {
    const int N = I.natoms, V = I.natoms + I.nnode;
    if (!(ff == ff) || isinf(ff)) {
        if (threadIdx.x == 0) B.status[s] = ST_NAN;
        return;
    }
    int4 xs = B.xstate[s];
    float4 st = B.fire[s];
    const float crit = (P.conv_mode == CONV_F2SUM) ? ff : fm2;
    bool reset_fire = false;
    if (xs.x == 0) {
        const bool conv = crit < P.Fconv2;
        if (conv || (P.xn_relax > 0 && xs.y >= P.xn_relax)) {
            if (conv) {
                const int m = xs.z;
                if (m < P.xkmax) {
                    if (threadIdx.x == 0) B.minE[(size_t)s * P.xkmax + m] = e;
                    if (B.minpos)
                        for (int v = threadIdx.x; v < V; v += blockDim.x)
                            B.minpos[(size_t)m * B.ntotV + I.v0 + v] = B.atoms[I.v0 + v];
                }
                xs.z++;
            }
            xs.x = 1;
            xs.y = 0;
        }
    } else if (xs.y >= P.xn_explore) {
        xs.x = 0;
        xs.y = 0;
        reset_fire = true;
    }
    __syncthreads();  // every thread has read xstate and the FIRE state
    if (step >= P.nstep_max) {
        if (threadIdx.x == 0) {
            B.status[s] = ST_MAXSTEP;
            B.xstate[s] = xs;
        }
        return;
    }
    const bool drive = xs.x == 1;
    float dt, cv, cfe;
    if (drive) {
        dt = P.fire.dt_max;
        cv = 1.f;
        cfe = 0.f;
    } else {
        FireState fs;
        fs.dt = reset_fire ? P.fire.dt_max : st.x;
        fs.damping = reset_fire ? P.fire.damp_max : st.y;
        fs.lastNeg = reset_fire ? 0 : (int)st.z;
        fire_update(P.fire, fs, vf, vv, ff, cv, cfe);
        dt = fs.dt;
        st = make_float4(fs.dt, fs.damping, (float)fs.lastNeg, 0.f);
    }
    if (threadIdx.x == 0) {
        B.fire[s] = st;
        xs.y++;
        B.xstate[s] = xs;
    }
    const float amp = drive ? sqrtf(2.f * BOLTZMANN_K * P.xT * P.xgamma / dt) : 0.f;
    for (int v = threadIdx.x; v < V; v += blockDim.x) {
        const bool isPi = v >= N;
        V3 p = xyz(B.atoms[I.v0 + v]);
        V3 f = xyz(B.aforce[I.v0 + v]);
        const V3 w0 = xyz(B.avel[I.v0 + v]);
        if (isPi) {
            f -= p * dot(p, f);
        } else if (drive) {
            f -= w0 * P.xgamma;
            f.x += amp * drive_xi(P.xseed, (unsigned)s, (unsigned)v, 0u, (unsigned)step);
            f.y += amp * drive_xi(P.xseed, (unsigned)s, (unsigned)v, 1u, (unsigned)step);
            f.z += amp * drive_xi(P.xseed, (unsigned)s, (unsigned)v, 2u, (unsigned)step);
        }
        V3 w = w0 * cv + f * (dt + cfe);
        if (isPi) w -= p * dot(p, w);
        p += w * dt;
        if (isPi) p = p * (1.f / sqrtf(dot(p, p)));
        B.avel[I.v0 + v] = make_float4(w.x, w.y, w.z, 0.f);
        B.atoms[I.v0 + v] = make_float4(p.x, p.y, p.z, 0.f);
    }
}

__global__ void __launch_bounds__(256) k_update(BatchDev B, RunParams P, int do_move, int step) {
    __shared__ float sh[8];
    __shared__ double shd[8];
    const int s = blockIdx.x;
    if (B.status[s] != ST_RUNNING) return;
    const SysInfo I = B.info[s];
    const int N = I.natoms, M = I.nnode, V = N + M;
    const float4* fng = B.fneigh;  // bkNeighs are absolute indices into fneigh
    float vf = 0.f, vv = 0.f, ff = 0.f, fm2 = 0.f;
    double e = 0.0;
    const int sflags = B.sysflags ? B.sysflags[s] : 0;  // restraints (bit 0), walls (bit 1)
    for (int v = threadIdx.x; v < V; v += blockDim.x) {
        const float4 a = B.aforce[I.v0 + v];
        V3 f = xyz(a);
        if (v < N) e += B.enb[I.a0 + v];
        if (v < N && sflags) {  // restraints and walls are part of the exact force
            const V3 p = xyz(B.atoms[I.v0 + v]);
            if (sflags & 1) f += restraint_force(p, B.constr[I.a0 + v], B.constrK[I.a0 + v]);
            if (sflags & 2) f += wall_force(p, B.bbox[3 * s], B.bbox[3 * s + 1], B.bbox[3 * s + 2]);
        }
        const int4 b = B.bkNeighs[I.v0 + v];
        if (b.x >= 0) { f += xyz(fng[b.x]);
        if (b.y >= 0) { f += xyz(fng[b.y]);
        if (b.z >= 0) { f += xyz(fng[b.z]);
        if (b.w >= 0) { f += xyz(fng[b.w]); } } } }
        if (v >= N) {
            const V3 p = xyz(B.atoms[I.v0 + v]);
            f -= p * dot(p, f);
        }
        const float f2 = dot(f, f);
        fm2 = fmaxf(fm2, f2);  // convergence always on the exact force
        B.aforce[I.v0 + v] = make_float4(f.x, f.y, f.z, 0.f);
        if (do_move && P.Flimit > 0.f && f2 > P.Flimit * P.Flimit) {
            f = f * (P.Flimit / sqrtf(f2));
            B.aforce[I.v0 + v] = make_float4(f.x, f.y, f.z, 0.f);
        }
        const V3 w = xyz(B.avel[I.v0 + v]);
        vf += dot(w, f);
        vv += dot(w, w);
        ff += dot(f, f);
    }
    for (int k = threadIdx.x; k < M; k += blockDim.x) e += (double)B.ebond[I.n0 + k];
    vf = blk_sum(vf, sh);
    vv = blk_sum(vv, sh);
    ff = blk_sum(ff, sh);
    fm2 = blk_max(fm2, sh);
    e = blk_sum_d(e, shd);
    if (threadIdx.x == 0) {
        B.Esys[s] = e;
        B.Fmax[s] = sqrtf(fm2);
        B.nsteps[s] = step;
    }
    if (!do_move) return;
    if (P.explore) {
        explore_move(B, P, s, I, step, vf, vv, ff, fm2, e);
        return;
    }
    int status = ST_RUNNING;
    const float crit = (P.conv_mode == CONV_F2SUM) ? ff : fm2;
    if (!(ff == ff) || isinf(ff)) status = ST_NAN;
    else if (crit < P.Fconv2) status = ST_CONVERGED;
    else if (step >= P.nstep_max) status = ST_MAXSTEP;
    if (status != ST_RUNNING) {
        if (threadIdx.x == 0) B.status[s] = status;
        return;
    }
    float4 st = B.fire[s];
    FireState fs;
    fs.dt = st.x;
    fs.damping = st.y;
    fs.lastNeg = (int)st.z;
    float cv, cfe;
    fire_update(P.fire, fs, vf, vv, ff, cv, cfe);
    __syncthreads();
    if (threadIdx.x == 0) B.fire[s] = make_float4(fs.dt, fs.damping, (float)fs.lastNeg, 0.f);
    const float dt = fs.dt;
    for (int v = threadIdx.x; v < V; v += blockDim.x) {
        const bool isPi = v >= N;
        V3 p = xyz(B.atoms[I.v0 + v]);
        V3 f = xyz(B.aforce[I.v0 + v]);
        if (isPi) f -= p * dot(p, f);
        V3 w = xyz(B.avel[I.v0 + v]) * cv + f * (dt + cfe);
        if (isPi) w -= p * dot(p, w);
        p += w * dt;
        if (isPi) p = p * (1.f / sqrtf(dot(p, p)));
        B.avel[I.v0 + v] = make_float4(w.x, w.y, w.z, 0.f);
        B.atoms[I.v0 + v] = make_float4(p.x, p.y, p.z, 0.f);
    }
}

// Reset per-system state before a split-path relaxation.
__global__ void k_reset_state(BatchDev B, RunParams P) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B.nsys) return;
    B.status[s] = ST_RUNNING;
    B.nsteps[s] = 0;
    B.fire[s] = make_float4(P.fire.dt_max, P.fire.damp_max, 0.f, 0.f);
}
// The same for the systems with only[s] != 0 (velocities zeroed too); the
// others keep their state, so the split kernels skip the finished ones.
__global__ void __launch_bounds__(256) k_reset_subset(BatchDev B, RunParams P, const int* only) {
    const int s = blockIdx.x;
    if (!only[s]) return;
    const SysInfo I = B.info[s];
    for (int v = threadIdx.x; v < I.natoms + I.nnode; v += blockDim.x) B.avel[I.v0 + v] = make_float4(0, 0, 0, 0);
    if (threadIdx.x == 0) {
        B.status[s] = ST_RUNNING;
        B.nsteps[s] = 0;
        B.fire[s] = make_float4(P.fire.dt_max, P.fire.damp_max, 0.f, 0.f);
    }
}
__global__ void k_zero_f4(float4* a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = make_float4(0, 0, 0, 0);
}


// Preconditioned force of every system from the exact forces in B.aforce
// (one 256-thread block per system; systems must have N <= 256 and M <= 256).
// gains: optional learned gains, per node at [n0 + k] of gnode and per
// system component at [s * PC_MAXC + c] of gcomp; nullptr = P.pc_gstar/grigid.
__global__ void __launch_bounds__(256) k_precond(BatchDev B, RunParams P, float4* out, const float* gnode,
                                                 const float* gcomp) {
    constexpr int BLOCK = 256;
    constexpr int NW = BLOCK / 64;
    __shared__ float4 s_pos[BLOCK];
    __shared__ float4 s_fnb[BLOCK];
    __shared__ float4 s_fself[BLOCK];
    __shared__ float4 s_fng[4 * BLOCK];
    __shared__ float s_cred[NW][PC_MAXC][PC_NSUM];
    __shared__ float s_csum[PC_MAXC][PC_NSUM];
    __shared__ float s_gain[BLOCK + PC_MAXC];
    const int s = blockIdx.x;
    const SysInfo I = B.info[s];
    const int N = I.natoms, M = I.nnode;
    if (N > BLOCK || M > BLOCK) return;  // uniform per block
    const int tid = threadIdx.x;
    const bool hasAtom = tid < N, hasNode = tid < M;
    if (hasAtom) s_pos[tid] = B.atoms[I.v0 + tid];
    const bool learned = gnode != nullptr;
    if (learned) {
        if (hasNode) s_gain[tid] = gnode[I.n0 + tid];
        if (tid < PC_MAXC) s_gain[M + tid] = gcomp[s * PC_MAXC + tid];
    }
    int4 ng = make_int4(-1, -1, -1, -1), bk0 = ng;
    if (hasNode) ng = B.neighs[I.a0 + tid];
    if (hasAtom) {
        const int4 b = B.bkNeighs[I.v0 + tid];
        const int base = I.n0 * 8;
        bk0 = make_int4(b.x >= 0 ? b.x - base : -1, b.y >= 0 ? b.y - base : -1, b.z >= 0 ? b.z - base : -1,
                        b.w >= 0 ? b.w - base : -1);
    }
    V3 f = hasAtom ? xyz(B.aforce[I.v0 + tid]) : zero3();
    const int myc = hasAtom ? B.comp[I.a0 + tid] : 0;
    __syncthreads();
    precondition<BLOCK>(P, s_pos, s_fnb, s_fself, s_fng, &s_cred[0][0][0], &s_csum[0][0], s_gain, learned, N, M,
                        B.ncomp[s], myc, hasAtom, hasNode, ng, bk0, f);
    if (hasAtom) out[I.v0 + tid] = make_float4(f.x, f.y, f.z, 0.f);
    for (int v = N + tid; v < N + M; v += BLOCK) out[I.v0 + v] = B.aforce[I.v0 + v];
}

}  // namespace fcd

extern "C" hipError_t fc_launch_split_step(hipStream_t st, const fcd::BatchDev* B, const fcd::RunParams* P,
                                           const int* node_sys, int ntot_nodes, int maxN, int do_move, int step) {
    dim3 g1((maxN + 63) / 64, B->nsys);
    hipLaunchKernelGGL(fcd::k_nonbond_grid, g1, dim3(64), 0, st, *B, *P);
    hipLaunchKernelGGL(fcd::k_bonded, dim3((ntot_nodes + 255) / 256), dim3(256), 0, st, *B, *P, node_sys, ntot_nodes);
    hipLaunchKernelGGL(fcd::k_update, dim3(B->nsys), dim3(256), 0, st, *B, *P, do_move, step);
    return hipGetLastError();
}

extern "C" hipError_t fc_launch_reset(hipStream_t st, const fcd::BatchDev* B, const fcd::RunParams* P) {
    hipLaunchKernelGGL(fcd::k_reset_state, dim3((B->nsys + 255) / 256), dim3(256), 0, st, *B, *P);
    return hipGetLastError();
}

extern "C" hipError_t fc_launch_reset_subset(hipStream_t st, const fcd::BatchDev* B, const fcd::RunParams* P,
                                             const int* only) {
    hipLaunchKernelGGL(fcd::k_reset_subset, dim3(B->nsys), dim3(256), 0, st, *B, *P, only);
    return hipGetLastError();
}

extern "C" hipError_t fc_launch_zero_f4(hipStream_t st, float4* a, int n) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fcd::k_zero_f4, dim3((n + 255) / 256), dim3(256), 0, st, a, n);
    return hipGetLastError();
}

extern "C" hipError_t fc_launch_precond(hipStream_t st, const fcd::BatchDev* B, const fcd::RunParams* P, float4* out,
                                        const float* gnode, const float* gcomp) {
    hipLaunchKernelGGL(fcd::k_precond, dim3(B->nsys), dim3(256), 0, st, *B, *P, out, gnode, gcomp);
    return hipGetLastError();
}

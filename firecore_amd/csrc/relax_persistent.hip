// Persistent batched FIRE relaxation on gfx950.
//
// One workgroup relaxes one system at a time, entirely on the device: it
// pulls a system index from a global work queue (systems pre-sorted by
// decreasing cost, so the largest start first), stages the system in LDS,
// iterates force evaluation + per-system FIRE until the system's own Fmax is
// below tolerance (or it hits the step limit / goes non-finite), writes the
// result back and pulls the next system.  Retirement is therefore per system
// with no host round trip and no grid-wide barrier.
//
// Per FIRE step (reference recipe MolWorld_sp3::run_no_omp,
// MolWorld_sp3.h:2291-2474):
//   A. forces: thread i -> non-bonded pairs (i, i+k mod N), k <= N/2, with the
//      reactions in per-wave LDS rows (positions/params from LDS, exclusions
//      from an LDS bit matrix) + B-spline GridFF sample through the atom's
//      window cache; thread k -> bonded terms of node k (recoils into LDS,
//      MMFFsp3_loc.h:250-575)
//   B. assemble: thread v -> own force + reactions + recoils via
//      back-neighbours, pi force projected perpendicular to the pi-vector
//      (assemble_atom, :998-1021); optional rigid-body preconditioner
//      (precond.h); block reduction of <v|f>, <v|v>, <f|f> and max|f|^2
//   C. per-system FIRE (DynamicOpt.cpp:250-267) and move (move_atom_FIRE,
//      MMFFsp3_loc.h:1264-1281)
// The learned form pauses the step loop every P steps to evaluate the
// proposal network (proposal.h) and restarts that step with the new gains.
#include "kernels.h"
#include "precond.h"
#include "proposal.h"

namespace fcd {

template <int BLOCK>
struct BlockRed {
    static constexpr int NW = BLOCK / 64;
};

// LDS layout (bytes), computed identically on the host (world.cpp).
//   pos   float4[nmax+mmax]     positions of atoms and pi-vectors
//   req   float4[nmax]          (R0, sqrt(E0), Q, H)
//   fnb   float4[nmax]          non-bonded + grid force accumulated by atom i's thread
//   facc  float4[nwave*frow]    per-wave reaction forces of the half-ring schedule:
//                              wave w's row holds partners 64w+1 .. 64w+63+nmax/2
//   fself float4[2*mmax]        bonded self forces of nodes and pi-vectors
//   fng   float4[8*mmax]        recoils (sigma rows, then pi rows)
//   mask  uint[wmax*nmax]       rotated exclusion rows, [word][atom]; only the
//                              words of offsets k <= nmax/2 (wmax = nmax/64 + 1)
//   learned gains only:
//   gain  float[mmax+PC_MAXC]   gains of the stars and the components
//   h     bf16 rows            network activations (one of the two buffers),
//                              aliased onto facc..mask, which is padded to fit
struct LdsLayout {
    int nmax, mmax, wmax, nwave, learned, h2lds;
    __host__ __device__ static int r16(int b) { return (b + 15) & ~15; }
    __host__ __device__ int vmax() const { return nmax + mmax; }
    __host__ __device__ int off_pos() const { return 0; }
    __host__ __device__ int off_req() const { return off_pos() + vmax() * 16; }
    __host__ __device__ int off_fnb() const { return off_req() + nmax * 16; }
    __host__ __device__ int off_facc() const { return off_fnb() + nmax * 16; }
    // entries of a wave's reaction row: lane l's partner at offset k lands at l+k-1
    __host__ __device__ int frow() const { return 64 + (nmax >> 1); }
    __host__ __device__ int off_fself() const { return off_facc() + nwave * frow() * 16; }
    __host__ __device__ int off_fng() const { return off_fself() + 2 * mmax * 16; }
    __host__ __device__ int off_mask() const { return off_fng() + 8 * mmax * 16; }
    __host__ __device__ int end_base() const { return off_mask() + nmax * wmax * 4; }
    __host__ __device__ int hbytes() const { return r16(nmax * MLP_HS * 2); }
    __host__ __device__ int off_gain() const {
        const int e = end_base();
        return r16((learned && off_facc() + hbytes() > e) ? off_facc() + hbytes() : e);
    }
    __host__ __device__ int off_hx() const { return off_gain() + r16((mmax + PC_MAXC) * 4); }
    // first activation buffer: in LDS over facc..mask, which is free during the
    // network evaluation (the exclusion rows are staged again after it) and is
    // padded to hold it (off_gain)
    __host__ __device__ int off_h2() const { return off_hx(); }
    __host__ __device__ int bytes() const {
        if (!learned) return end_base();
        return off_h2() + (h2lds ? hbytes() : 0);
    }
};

// Work queues: class c owns order[qoff[c] .. qoff[c]+qcnt[c]) and counter
// queue[c].  A workgroup of class `cls` drains its own queue, then steals from
// the queues of the smaller classes (its LDS layout and width fit them), so no
// launch idles while another still has systems.
struct Queues {
    const int* order;
    int qoff[4];
    int qcnt[4];
    int* counter;
    int lowest;  // smallest class index a workgroup may steal from (its own index: no stealing)
    unsigned short* hslot;  // learned gains: global activation rows, one slot per workgroup of this launch
};

// minimum resident waves per SIMD the register allocation must allow
#ifndef FC_WAVES_PER_SIMD
#define FC_WAVES_PER_SIMD 3
#endif
#ifndef FC_WAVES_PER_SIMD_LEARNED
#define FC_WAVES_PER_SIMD_LEARNED FC_WAVES_PER_SIMD
#endif
template <int BLOCK, bool HB, int PC>
__global__ void __launch_bounds__(BLOCK, PC == 2 ? FC_WAVES_PER_SIMD_LEARNED : FC_WAVES_PER_SIMD)
relax_persistent_kernel(BatchDev B, RunParams P, Queues Q, int cls,
                                                                 LdsLayout L) {
    extern __shared__ float4 lds_dyn[];
    char* lds = reinterpret_cast<char*>(lds_dyn);
    float4* s_pos = reinterpret_cast<float4*>(lds + L.off_pos());
    float4* s_req = reinterpret_cast<float4*>(lds + L.off_req());
    float4* s_fnb = reinterpret_cast<float4*>(lds + L.off_fnb());
    float4* s_fself = reinterpret_cast<float4*>(lds + L.off_fself());
    float4* s_fng = reinterpret_cast<float4*>(lds + L.off_fng());
    unsigned* s_mask = reinterpret_cast<unsigned*>(lds + L.off_mask());
    float4* s_facc = reinterpret_cast<float4*>(lds + L.off_facc());
    float* s_gain = reinterpret_cast<float*>(lds + L.off_gain());
    // activation buffers of the in-kernel network: LDS where the layout allows,
    // else the two halves of this workgroup's global slot
    unsigned short* s_h = reinterpret_cast<unsigned short*>(lds + L.off_facc());
    unsigned short* g_h = L.h2lds ? reinterpret_cast<unsigned short*>(lds + L.off_h2())
                                  : (Q.hslot ? Q.hslot + (size_t)blockIdx.x * (L.hbytes() / 2) : nullptr);
    const bool learned = PC >= 2 && L.learned && P.pc_mode == 2;
    const bool inkernel = learned && P.chunk == 0;  // network evaluated inside this kernel
    constexpr int NW = BLOCK / 64;
    __shared__ float s_red[4][2 * NW];
    __shared__ double s_redd[NW];
    __shared__ int s_q;
    // preconditioner: per-wave partial sums and per-component totals (PC_NSUM values each)
    __shared__ float s_cred[PC ? NW : 1][PC ? PC_MAXC : 1][PC_NSUM];
    __shared__ float s_csum[PC ? PC_MAXC : 1][PC_NSUM];
    __shared__ float s_smax[NW];
    __shared__ float s_st_best;       // FIRE fallback of the learned form (below)
    __shared__ int s_st_cnt, s_st_plain;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wid = tid >> 6;

    int cc = cls;
    for (;;) {
        if (tid == 0) {
            int q = -1;
            while (cc >= Q.lowest) {
                const int k = atomicAdd(Q.counter + cc, 1);
                if (k < Q.qcnt[cc]) {
                    q = Q.order[Q.qoff[cc] + k];
                    if (P.chunk > 0 && B.status[q] != ST_RUNNING) continue;  // finished in an earlier chunk
                    break;
                }
                cc--;
            }
            s_q = q;
        }
        __syncthreads();
        const int q = s_q;
        __syncthreads();
        if (q < 0) break;
        const int s = q;
        if (B.systime && tid == 0 && P.chunk == 0) B.systime[2 * s] = (long long)wall_clock64();
        const SysInfo I = B.info[s];
        const int N = I.natoms, M = I.nnode, V = N + M;
        const int sflags = B.sysflags ? B.sysflags[s] : 0;  // restraints (bit 0), walls (bit 1)
        const int W = (N + 31) >> 5;         // exclusion words per atom in memory
        const int Wh = ((N >> 1) >> 5) + 1;  // words of the offsets k <= N/2 the half ring reads
        // exclusion rows transposed to [word][atom] so that lanes read consecutive banks
        auto stage_mask = [&] {
            for (int w = tid; w < N * Wh; w += BLOCK) {
                const int a = w / Wh, wd = w - a * Wh;
                s_mask[wd * N + a] = B.excl_rot[I.mask0 + a * W + wd];
            }
        };

        // ---- stage the system in LDS
        for (int v = tid; v < V; v += BLOCK) s_pos[v] = B.atoms[I.v0 + v];
        for (int i = tid; i < N; i += BLOCK) s_req[i] = B.REQs[I.a0 + i];
        stage_mask();

        const bool hasAtom = tid < N;
        const bool hasNode = tid < M;
        const int nc = PC ? B.ncomp[s] : 0;
        V3 vel[2];
#pragma unroll
        for (int r = 0; r < 2; r++) vel[r] = zero3();
        FireState fs;
        fs.dt = P.fire.dt_max;
        fs.damping = P.fire.damp_max;
        fs.lastNeg = 0;
        int step = 0;
        int gkey = -1;    // grid window held in this atom's cache slot (none yet)
        if (P.chunk > 0) {  // resume the suspended state of this system
            const float4 fst = B.fire[s];
            fs.dt = fst.x;
            fs.damping = fst.y;
            fs.lastNeg = (int)fst.z;
            step = B.nsteps[s];
#pragma unroll
            for (int r = 0; r < 2; r++) {
                const int v = tid + r * BLOCK;
                if (v < V) vel[r] = xyz(B.avel[I.v0 + v]);
            }
        }
        // FIRE fallback of the learned form: the best block max|F|^2 seen at a
        // proposal point after step 0, the proposal points since it was last
        // improved, and whether the system has fallen back to plain FIRE (zero
        // gains, no more network evaluations) after P.pc_stall such points.
        // Chunks of the chunked form carry it in B.stall.
        // The state lives in LDS, written by thread 0 between barriers only.
        if (learned) {
            const bool resume = P.chunk > 0 && step > 0;
            const float4 t = resume ? B.stall[s] : make_float4(INFINITY, 0.f, 0.f, 0.f);
            if (tid == 0) {
                s_st_best = t.x;
                s_st_cnt = (int)t.y;
                s_st_plain = t.z != 0.f ? 1 : 0;
            }
            if (!inkernel) {  // chunked: gains from the network kernel between chunks
                const bool plain = t.z != 0.f;
                for (int k = tid; k < M; k += BLOCK) s_gain[k] = plain ? 0.f : B.gnode[I.n0 + k];
                if (tid < PC_MAXC) s_gain[M + tid] = plain ? 0.f : B.gcomp[s * PC_MAXC + tid];
            }
        }
        // called by every thread of the block with its own max|F|^2 at a proposal point
        auto stall_check = [&](float fm2t) {
            fm2t = wave_max(fm2t);
            if (lane == 0) s_smax[wid] = fm2t;
            __syncthreads();
            if (tid == 0) {
                float m = s_smax[0];
                for (int w = 1; w < NW; w++) m = fmaxf(m, s_smax[w]);
                if (m < s_st_best) {
                    s_st_best = m;
                    s_st_cnt = 0;
                } else if (++s_st_cnt >= P.pc_stall) {
                    s_st_plain = 1;
                }
            }
            __syncthreads();
            if (s_st_plain)
                for (int k = tid; k < M + PC_MAXC; k += BLOCK) s_gain[k] = 0.f;
            __syncthreads();
        };
        const int step_end = (P.chunk > 0) ? step + P.chunk : 0x7fffffff;
        bool suspended = false;
        __syncthreads();

        int status = ST_RUNNING;
        float fmax2_last = 0.f;
        V3 fv[2];
        // The learned form evaluates the network between segments of the step
        // loop (one segment per proposal period): the step loop breaks out at a
        // proposal step once its exact forces are assembled, the network runs
        // with the forces and velocities parked in memory, and the segment
        // resumes that step at the preconditioner.  Nothing of the step loop is
        // live across the network, so neither part adds to the other's registers.
        // gains of the current step proposed: resume it after its force assembly.
        // A chunk of the chunked learned form resumes the step it was suspended
        // at, whose exact forces it parked (the first chunk computes its own)
        bool fresh = learned && !inkernel && step > 0;
        for (;;) {
            bool propose_now = false;
            // ---- per-thread state of this segment: atom tid, node tid, vecs tid and tid+BLOCK
            const float4 plq = hasAtom ? B.PLQs[I.a0 + tid] : make_float4(0, 0, 0, 0);
            const int myc = (PC && hasAtom) ? B.comp[I.a0 + tid] : 0;
            int4 ng = make_int4(-1, -1, -1, -1);
            float4 apar = make_float4(0, 0, 0, 0), bL = apar, bK = apar, ksp = apar, kpp = apar;
            if (hasNode) {
                ng = B.neighs[I.a0 + tid];
                apar = B.apars[I.n0 + tid];
                bL = B.bLs[I.n0 + tid];
                bK = B.bKs[I.n0 + tid];
                ksp = B.Ksp[I.n0 + tid];
                kpp = B.Kpp[I.n0 + tid];
            }
            int4 bk[2];
#pragma unroll
            for (int r = 0; r < 2; r++) {
                const int v = tid + r * BLOCK;
                bk[r] = make_int4(-1, -1, -1, -1);
                if (v < V) {
                    // back-neighbours as local indices into s_fng (sigma rows [0,4M), pi rows [4M,8M))
                    int4 b = B.bkNeighs[I.v0 + v];
                    const int base = I.n0 * 8;
                    bk[r].x = b.x >= 0 ? b.x - base : -1;
                    bk[r].y = b.y >= 0 ? b.y - base : -1;
                    bk[r].z = b.z >= 0 ? b.z - base : -1;
                    bk[r].w = b.w >= 0 ? b.w - base : -1;
                }
            }
            for (;;) {
                if (B.snap && step % B.snap_every == 0 && step / B.snap_every < B.nsnap) {
                    float4* dst = B.snap + (size_t)(step / B.snap_every) * B.ntotV + I.v0;
                    for (int v = tid; v < V; v += BLOCK) dst[v] = s_pos[v];
                }
                float vf = 0.f, vv = 0.f, ff = 0.f, fm2 = 0.f;
                if (!fresh) {
                    // ================= A: forces (no energies inside the step loop)
                    // Half-ring pair schedule (Newton's third law): the thread of atom i
                    // takes the partners j = (i+k) mod N, k = 1..N/2, so every unordered
                    // pair is computed once.  The partners of a wave at offset k are the
                    // 64 consecutive atoms 64w+k .. 64w+63+k (mod N), a window that slides
                    // by one atom per offset.  The reaction on each partner is accumulated
                    // in registers and moves down one lane per offset with the window
                    // (DPP wave shift; the atom entering at lane 63 starts from zero); the
                    // reaction leaving at lane 0 is written to the wave's reaction row.
                    // Entry e of wave w's row holds the reaction on atom (64w+1+e) mod N:
                    // entries 0..N/2-2 are written as they leave, the rest by all 64 lanes
                    // after the last offset.  Every entry is written once and read after
                    // the block barrier: no clearing, no atomics, no lane ordering.
                    const int nwAct = (N + 63) >> 6;
                    const int Kl = N >> 1;          // last offset of the half ring
                    const int nrow = 63 + Kl;       // entries of a wave's reaction row
                    const int wu = __builtin_amdgcn_readfirstlane(wid);
                    V3 f = zero3();
                    // lanes without an atom carry the window only: their own atom sits far
                    // away with zero parameters, so each of their pair forces is exactly 0
                    const V3 pi = hasAtom ? xyz(s_pos[tid]) : mk(1e4f, 1e4f, 1e4f);
                    const float4 ri = hasAtom ? s_req[tid] : make_float4(0.f, 0.f, 0.f, 0.f);
                    const float kQi = COULOMB_K * ri.z;
                    if (P.bNonBonded && wu < nwAct && N >= 2) {
                        float4* facc = s_facc + wu * L.frow();
                        const int ic = hasAtom ? tid : 0;
                        unsigned j = (unsigned)((wu * 64 + lane + 1) % N);  // partner at offset 1
                        V3 acc = zero3();
                        // one offset: the pair (i, j) with partner rows (pj, rj), excluded when
                        // bit k of the offset's mask word is set; reads the next offset's
                        // partner rows into (pn, rn) ahead of the arithmetic
                        auto pair_step = [&](unsigned mword, int k, const float4& pj, const float4& rj, float4& pn,
                                             float4& rn) {
                            j = min(j + 1, j + 1 - (unsigned)N);  // (j+1) mod N
                            pn = s_pos[j];
                            rn = s_req[j];
                            const V3 d = mk(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z);
                            float fr, ep;
                            lj_pair<HB>(d, ri.x + rj.x, ri.y * rj.y, kQi * rj.z, ri.w * rj.w, fr, ep);
                            fr = ((mword >> (k & 31)) & 1u) ? 0.f : fr;
                            f.x = __builtin_fmaf(-d.x, fr, f.x);
                            f.y = __builtin_fmaf(-d.y, fr, f.y);
                            f.z = __builtin_fmaf(-d.z, fr, f.z);
                            acc.x = __builtin_fmaf(d.x, fr, acc.x);
                            acc.y = __builtin_fmaf(d.y, fr, acc.y);
                            acc.z = __builtin_fmaf(d.z, fr, acc.z);
                        };
                        // the window moves on: the reaction of lane 0's partner leaves
                        auto shift = [&](int k) {
                            if (lane == 0) facc[k - 1] = make_float4(acc.x, acc.y, acc.z, 0.f);
                            acc.x = dpp_shl1_zero(acc.x);
                            acc.y = dpp_shl1_zero(acc.y);
                            acc.z = dpp_shl1_zero(acc.z);
                        };
                        float4 pA = s_pos[j], rA = s_req[j], pB, rB;
                        // offsets 1..Kl-1, two per iteration (partner rows alternate A/B)
                        for (int w = 0; (w << 5) < Kl; w++) {
                            const unsigned m = s_mask[w * N + ic];
                            const int k0 = max(1, w << 5);
                            const int k1 = min(Kl - 1, (w << 5) + 31);
                            int k = k0;
                            for (; k < k1; k += 2) {
                                pair_step(m, k, pA, rA, pB, rB);
                                shift(k);
                                pair_step(m, k + 1, pB, rB, pA, rA);
                                shift(k + 1);
                            }
                            if (k == k1) {
                                pair_step(m, k, pA, rA, pB, rB);
                                shift(k);
                                pA = pB;
                                rA = rB;
                            }
                        }
                        // last offset Kl: for even N the pair (i, i+N/2) is taken by the lower half only
                        {
                            unsigned m = s_mask[(Kl >> 5) * N + ic];
                            if (!(N & 1) && tid >= Kl) m |= 1u << (Kl & 31);
                            pair_step(m, Kl, pA, rA, pB, rB);
                        }
                        facc[lane + Kl - 1] = make_float4(acc.x, acc.y, acc.z, 0.f);
                    }
                    if (hasAtom) {
                        const int i = tid;
                        if (P.bNonBonded) {
                            // weighted corrections (ring topologies; empty for trees)
                            for (int c = 0; c < I.ncorr; c++) {
                                const PairCorrDev pc = B.corr[I.corr0 + c];
                                if (pc.i != i) continue;
                                const float4 pj = s_pos[pc.j];
                                const float4 rj = s_req[pc.j];
                                const V3 d = mk(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z);
                                float fr, ep;
                                lj_pair<HB>(d, ri.x + rj.x, ri.y * rj.y, kQi * rj.z, ri.w * rj.w, fr, ep);
                                f -= d * (fr * pc.w);
                            }
                        }
                        if (P.bGridFF) {
                            grid_sample_cached(P.grid, pi, mk(plq.x, plq.y, plq.z), B.gridcache + I.a0 + i,
                                               B.gridcache_stride, gkey, f);
                        }
                        s_fnb[i] = make_float4(f.x, f.y, f.z, 0.f);
                    }
                    if (hasNode) {
                        V3 fbs[4], fps[4], fa, fpi;
                        auto pos = [&](int v) { return xyz(s_pos[v]); };
                        eval_node(tid, N, M, pos, ng, apar, bL, bK, ksp, kpp, P.sw, fbs, fps, fa, fpi);
#pragma unroll
                        for (int k = 0; k < 4; k++) {
                            s_fng[tid * 4 + k] = make_float4(fbs[k].x, fbs[k].y, fbs[k].z, 0.f);
                            s_fng[4 * M + tid * 4 + k] = make_float4(fps[k].x, fps[k].y, fps[k].z, 0.f);
                        }
                        s_fself[tid] = make_float4(fa.x, fa.y, fa.z, 0.f);
                        s_fself[M + tid] = make_float4(fpi.x, fpi.y, fpi.z, 0.f);
                    }
                    __syncthreads();

                    // ================= B: assemble + reductions
#pragma unroll
                    for (int r = 0; r < 2; r++) {
                        const int v = tid + r * BLOCK;
                        fv[r] = zero3();
                        if (v < V) {
                            V3 f = zero3();
                            const bool isPi = v >= N;
                            if (!isPi) {
                                f = xyz(s_fnb[v]);
                                if (P.bNonBonded && N >= 2)
                                    for (int w = 0; w < nwAct; w++) {  // atom v: entries (v-64w-1) mod N + multiples of N
                                        int e = (v - 64 * w - 1) % N;
                                        e = (e < 0) ? e + N : e;
                                        const float4* row = s_facc + w * L.frow();
                                        for (; e < nrow; e += N) f += xyz(row[e]);
                                    }
                                if (v < M) f += xyz(s_fself[v]);
                                if (sflags) {  // restraints and walls are part of the exact force
                                    const V3 p = xyz(s_pos[v]);
                                    if (sflags & 1) f += restraint_force(p, B.constr[I.a0 + v], B.constrK[I.a0 + v]);
                                    if (sflags & 2) f += wall_force(p, B.bbox[3 * s], B.bbox[3 * s + 1], B.bbox[3 * s + 2]);
                                }
                            } else {
                                f = xyz(s_fself[M + (v - N)]);
                            }
                            const int b4[4] = {bk[r].x, bk[r].y, bk[r].z, bk[r].w};
#pragma unroll
                            for (int k = 0; k < 4; k++) {
                                if (b4[k] < 0) break;
                                f += xyz(s_fng[b4[k]]);
                            }
                            if (isPi) {
                                const V3 p = xyz(s_pos[v]);
                                f -= p * dot(p, f);
                            }
                            fm2 = fmaxf(fm2, dot(f, f));  // convergence always on the exact force
                            fv[r] = f;
                        }
                    }
                } else {
                    // restart of a proposal step: its exact forces were parked before the network
#pragma unroll
                    for (int r = 0; r < 2; r++) {
                        const int v = tid + r * BLOCK;
                        fv[r] = zero3();
                        if (v < V) {
                            fv[r] = xyz(B.aforce[I.v0 + v]);
                            fm2 = fmaxf(fm2, dot(fv[r], fv[r]));
                        }
                    }
                }
                // a proposal point after step 0 (in-kernel: before the network runs;
                // chunked: at the start of a chunk, after the network kernel)
                if (learned && P.pc_stall > 0 && step > 0 && step % P.pc_every == 0 && (inkernel ? !fresh : fresh) &&
                    !s_st_plain)
                    stall_check(fm2);
                if (!fresh && step >= step_end) {
                    // chunk done: suspend with the exact forces of the current
                    // geometry (written back below), so that the network between
                    // chunks sees positions and forces of the same step
                    suspended = true;
                    break;
                }
                if (PC == 2 && inkernel && !fresh && step % P.pc_every == 0 && !s_st_plain) {  // uniform across the block
                    if (hasAtom) s_fnb[tid] = make_float4(fv[0].x, fv[0].y, fv[0].z, 0.f);
#pragma unroll
                    for (int r = 0; r < 2; r++) {
                        const int v = tid + r * BLOCK;
                        if (v < V) B.aforce[I.v0 + v] = make_float4(fv[r].x, fv[r].y, fv[r].z, 0.f);
                    }
                    propose_now = true;
                    break;
                }
                if (PC)
                    precondition<BLOCK>(P, s_pos, s_fnb, s_fself, s_fng, &s_cred[0][0][0], &s_csum[0][0], s_gain,
                                        learned, N, M, nc, myc, hasAtom, hasNode, ng, bk[0], fv[0]);
                // FIRE sums: totals per chunk of 64 consecutive vecs (vec v = tid + r*BLOCK
                // lies in chunk wid + r*NW), added in chunk order, so that the result does
                // not depend on the workgroup width a system is relaxed with
#pragma unroll
                for (int r = 0; r < 2; r++) {
                    const int v = tid + r * BLOCK;
                    float pvf = 0.f, pvv = 0.f, pff = 0.f;
                    if (v < V) {
                        V3 f = fv[r];
                        const float f2 = dot(f, f);
                        if (P.Flimit > 0.f && f2 > P.Flimit * P.Flimit) f = f * (P.Flimit / sqrtf(f2));
                        fv[r] = f;
                        pvf = dot(vel[r], f);
                        pvv = dot(vel[r], vel[r]);
                        pff = dot(f, f);
                    }
                    const float t = wave_sum4(pvf, pvv, pff, 0.f);  // lane l: total of value l >> 4
                    if ((lane & 15) == 0 && lane < 48) s_red[lane >> 4][wid + r * NW] = t;
                }
                fm2 = wave_max(fm2);
                if (lane == 0) s_red[3][wid] = fm2;
                __syncthreads();
                vf = 0.f;
                vv = 0.f;
                ff = 0.f;
#pragma unroll
                for (int c = 0; c < 2 * NW; c++) {
                    vf += s_red[0][c];
                    vv += s_red[1][c];
                    ff += s_red[2][c];
                }
                fm2 = s_red[3][0];
#pragma unroll
                for (int w = 1; w < NW; w++) fm2 = fmaxf(fm2, s_red[3][w]);
                fmax2_last = fm2;
                const float crit = (P.conv_mode == CONV_F2SUM) ? ff : fm2;
                if (!(ff == ff) || isinf(ff)) status = ST_NAN;
                else if (crit < P.Fconv2) status = ST_CONVERGED;
                else if (step >= P.nstep_max) status = ST_MAXSTEP;
                if (status != ST_RUNNING) {
                    // energy of this geometry: precise 1/r, FP64 accumulation and reduction
                    double epart = 0.0;
                    if (hasAtom && P.bNonBonded) {
                        const int i = tid;
                        const V3 pi = xyz(s_pos[i]);
                        const float4 ri = s_req[i];
                        const float kQi = COULOMB_K * ri.z;
                        double e = 0.0;  // each unordered pair once (half ring)
                        const int Kh = N >> 1;
                        for (int k = 1; k <= Kh; k++) {
                            if (!(N & 1) && k == Kh && i >= Kh) break;
                            if ((s_mask[(k >> 5) * N + i] >> (k & 31)) & 1u) continue;
                            int j = i + k;
                            j = (j >= N) ? j - N : j;
                            const float4 pj = s_pos[j];
                            const float4 rj = s_req[j];
                            float fr, ep;
                            lj_pair<true, true>(mk(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z), ri.x + rj.x, ri.y * rj.y,
                                                kQi * rj.z, ri.w * rj.w, fr, ep);
                            e += (double)ep;
                        }
                        double ec = 0.0;  // directed corrections carry half of the pair energy each
                        for (int c = 0; c < I.ncorr; c++) {
                            const PairCorrDev pc = B.corr[I.corr0 + c];
                            if (pc.i != i) continue;
                            const float4 pj = s_pos[pc.j];
                            const float4 rj = s_req[pc.j];
                            float fr, ep;
                            lj_pair<true, true>(mk(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z), ri.x + rj.x, ri.y * rj.y,
                                                kQi * rj.z, ri.w * rj.w, fr, ep);
                            ec += (double)ep * pc.w;
                        }
                        epart += e + 0.5 * ec;
                    }
                    if (hasAtom && P.bGridFF) {
                        V3 fd = zero3();
                        epart += (double)grid_sample(P.grid, xyz(s_pos[tid]), mk(plq.x, plq.y, plq.z), fd);
                    }
                    if (hasNode) {
                        V3 fbs[4], fps[4], fa, fpi;
                        auto pos = [&](int v) { return xyz(s_pos[v]); };
                        epart += (double)eval_node(tid, N, M, pos, ng, apar, bL, bK, ksp, kpp, P.sw, fbs, fps, fa, fpi);
                    }
                    double ed = wave_sum_d(epart);
                    if (lane == 0) s_redd[wid] = ed;
                    __syncthreads();
                    if (tid == 0) {
                        double E = 0.0;
                        for (int w = 0; w < NW; w++) E += s_redd[w];
                        B.Esys[s] = E;
                        B.nsteps[s] = step;
                        B.status[s] = status;
                        B.Fmax[s] = sqrtf(fm2);
                    }
                    break;
                }

                // ================= C: FIRE + move
                float cv, cfe;
                fire_update(P.fire, fs, vf, vv, ff, cv, cfe);
                const float dt = fs.dt;
#pragma unroll
                for (int r = 0; r < 2; r++) {
                    const int v = tid + r * BLOCK;
                    if (v < V) {
                        const bool isPi = v >= N;
                        V3 p = xyz(s_pos[v]);
                        V3 f = fv[r];
                        if (isPi) f -= p * dot(p, f);
                        V3 w = vel[r] * cv + f * (dt + cfe);
                        if (isPi) w -= p * dot(p, w);
                        p += w * dt;
                        if (isPi) p = p * (1.f / sqrtf(dot(p, p)));
                        vel[r] = w;
                        s_pos[v] = make_float4(p.x, p.y, p.z, 0.f);
                    }
                }
                step++;
                fresh = false;
                __syncthreads();
            }
            if (!propose_now) break;
            // ---- network evaluation between segments (learned form)
#pragma unroll
            for (int r = 0; r < 2; r++) {
                const int v = tid + r * BLOCK;
                if (v < V) B.avel[I.v0 + v] = make_float4(vel[r].x, vel[r].y, vel[r].z, 0.f);
            }
            __syncthreads();
            propose_gains<BLOCK>(P, B.neighs + I.a0, B.REQs + I.a0, B.elem + I.a0, s_pos, s_fnb, s_req, s_h, g_h,
                                 s_gain, &s_cred[0][0][0], N, M, nc, hasAtom ? B.comp[I.a0 + tid] : 0,
                                 hasAtom ? B.PLQs[I.a0 + tid] : make_float4(0, 0, 0, 0));
            stage_mask();  // the exclusion rows were overwritten by the activations
#pragma unroll
            for (int r = 0; r < 2; r++) {
                const int v = tid + r * BLOCK;
                if (v < V) vel[r] = xyz(B.avel[I.v0 + v]);
            }
            fresh = true;
            __syncthreads();
        }
        (void)fmax2_last;
        if (suspended && tid == 0) {
            B.fire[s] = make_float4(fs.dt, fs.damping, (float)fs.lastNeg, 0.f);
            B.nsteps[s] = step;
        }
        if (learned && tid == 0) B.stall[s] = make_float4(s_st_best, (float)s_st_cnt, s_st_plain ? 1.f : 0.f, 0.f);
        // ---- write back the final state of this system
#pragma unroll
        for (int r = 0; r < 2; r++) {
            const int v = tid + r * BLOCK;
            if (v < V) {
                B.atoms[I.v0 + v] = s_pos[v];
                B.avel[I.v0 + v] = make_float4(vel[r].x, vel[r].y, vel[r].z, 0.f);
                // preconditioned forms keep the exact atom forces in s_fnb; a suspended
                // chunk holds them, assembled and not yet preconditioned, in fv
                const V3 fo = (PC && v < N && !suspended) ? xyz(s_fnb[v]) : fv[r];
                B.aforce[I.v0 + v] = make_float4(fo.x, fo.y, fo.z, 0.f);
            }
        }
        if (B.systime && tid == 0 && P.chunk == 0) B.systime[2 * s + 1] = (long long)wall_clock64();
        __syncthreads();
    }
}

#define FC_INST(BL, HB, PCF) \
    template __global__ void relax_persistent_kernel<BL, HB, PCF>(BatchDev, RunParams, Queues, int, LdsLayout);
#define FC_INST2(BL)      \
    FC_INST(BL, false, 0) \
    FC_INST(BL, true, 0)  \
    FC_INST(BL, false, 1) \
    FC_INST(BL, true, 1)  \
    FC_INST(BL, false, 2) \
    FC_INST(BL, true, 2)  \
    FC_INST(BL, false, 3) \
    FC_INST(BL, true, 3)
FC_INST2(64)
FC_INST2(128)
FC_INST2(192)
FC_INST2(256)
#undef FC_INST2
#undef FC_INST

template <bool HB, int PCF>
static const void* kernel_ptr(int block) {
    switch (block) {
        case 64: return (const void*)&relax_persistent_kernel<64, HB, PCF>;
        case 128: return (const void*)&relax_persistent_kernel<128, HB, PCF>;
        case 192: return (const void*)&relax_persistent_kernel<192, HB, PCF>;
        case 256: return (const void*)&relax_persistent_kernel<256, HB, PCF>;
        default: return nullptr;
    }
}
// pc: preconditioner mode 0 (off), 1 (constant gains), 2 (learned gains, network
// inside the kernel), 3 (learned gains read from memory: chunked execution)
static const void* kernel_ptr(int block, int hb, int pc) {
    switch (pc) {
        case 1: return hb ? kernel_ptr<true, 1>(block) : kernel_ptr<false, 1>(block);
        case 2: return hb ? kernel_ptr<true, 2>(block) : kernel_ptr<false, 2>(block);
        case 3: return hb ? kernel_ptr<true, 3>(block) : kernel_ptr<false, 3>(block);
        default: return hb ? kernel_ptr<true, 0>(block) : kernel_ptr<false, 0>(block);
    }
}

}  // namespace fcd

// Host launcher (C linkage inside the library).  qinfo = {off[4], cnt[4]}.
extern "C" hipError_t fc_launch_relax_persistent(int block, int hb, int ngroups, size_t lds_bytes, hipStream_t stream,
                                                 const fcd::BatchDev* B, const fcd::RunParams* P, const int* order,
                                                 const int* qinfo, int* counters, int cls, int nmax, int mmax,
                                                 int wmax, int steal, unsigned short* hslot, int h2lds) {
    fcd::LdsLayout L{nmax, mmax, wmax, block / 64, P->pc_mode == 2 ? 1 : 0, h2lds};
    fcd::Queues Q;
    Q.order = order;
    for (int c = 0; c < 4; c++) {
        Q.qoff[c] = qinfo[c];
        Q.qcnt[c] = qinfo[4 + c];
    }
    Q.counter = counters;
    Q.lowest = steal ? 0 : cls;
    Q.hslot = hslot;
    const void* k = fcd::kernel_ptr(block, hb, (P->pc_mode == 2 && P->chunk > 0) ? 3 : P->pc_mode);
    if (!k) return hipErrorInvalidValue;
    void* args[] = {(void*)B, (void*)P, (void*)&Q, (void*)&cls, (void*)&L};
    return hipLaunchKernel(k, dim3(ngroups), dim3(block), args, lds_bytes, stream);
}

// Scratch of the allocation-free relaxation (World::relax_noalloc): the queue
// counters, the queue order, the activation slots of the in-kernel network and
// the GridFF window cache of one relaxation, in one file-scope device array of
// FC_NOALLOC_SCRATCH_BYTES.  Each device holds its own copy at its own address
// until the process ends; host code uses the address returned here, never the
// array's name.
static __device__ __align__(256) unsigned char scratch[FC_NOALLOC_SCRATCH_BYTES];

// Device address of this device's scratch array (looked up once per device,
// with that device current).
extern "C" hipError_t fc_noalloc_scratch(void** ptr) {
    constexpr int kMaxDev = 64;
    static void* dev_ptr[kMaxDev] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDev) return hipErrorInvalidDevice;
    if (!dev_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        dev_ptr[dev] = p;
    }
    *ptr = dev_ptr[dev];
    return hipSuccess;
}

extern "C" int fc_lds_bytes(int nmax, int mmax, int wmax, int block, int learned, int h2lds) {
    fcd::LdsLayout L{nmax, mmax, wmax, block / 64, learned, h2lds};
    return L.bytes();
}

// bf16 elements of one workgroup's global activation slot (the second buffer)
extern "C" int fc_hslot_elems(int nmax) {
    fcd::LdsLayout L{nmax, 1, 1, 1, 1, 0};
    return L.hbytes() / 2;
}

extern "C" hipError_t fc_persistent_occupancy(int block, int hb, int pc, size_t lds, int* nblocks) {
    const void* k = fcd::kernel_ptr(block, hb, pc);
    if (!k) return hipErrorInvalidValue;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(nblocks, k, block, lds);
}

// Kernel-side descriptors shared by the host world and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "ff_device.h"

// Size of the device scratch array of the allocation-free relaxation (queue
// counters, queue order, network activation slots, GridFF window cache).
#define FC_NOALLOC_SCRATCH_BYTES (64u << 20)

namespace fcd {

// One relaxation system in the CSR-packed batch.
struct SysInfo {
    int natoms, nnode;
    int a0;     // first atom   (REQs, neighs, PLQs, excl rows)
    int n0;     // first node   (apars, bLs, bKs, Ksp, Kpp; fneigh block n0*8)
    int v0;     // first vec    (atoms, avel, aforce, bkNeighs)
    int mask0;  // first word of the exclusion bit-matrix (natoms rows of W words)
    int corr0, ncorr;  // weighted pair corrections
    int sh0, npbc;     // periodic molecule: its image shifts pbcshifts[sh0 .. sh0+npbc); npbc = 0 in vacuum
};

// Weighted pair correction of the non-bonded sum for atom i against atom j.
// path < 0: at the nearest image (vacuum rings); path = c*16 + ki*4 + kj: a
// 1-3 pair of a periodic molecule through node c, neighbour slots ki (atom i)
// and kj (atom j), at the image neighCell[c][kj] - neighCell[c][ki].
struct PairCorrDev {
    int i, j;
    float w;
    int path;
};

// Status codes written per system.
enum { ST_RUNNING = 0, ST_CONVERGED = 1, ST_MAXSTEP = 2, ST_NAN = 3 };

// Convergence criterion.
enum { CONV_FMAX = 0, CONV_F2SUM = 1 };

struct BatchDev {
    int nsys;
    const SysInfo* info;
    float4* atoms;      // [sum V] positions (atoms, then pi-vectors)
    float4* avel;       // [sum V]
    float4* aforce;     // [sum V]
    const float4* REQs; // [sum N] (R0, sqrt(E0), Q, H)
    const float4* PLQs; // [sum N] (cP, cL, Q, cH)
    const int4* neighs; // [sum N] local indices
    const int4* bkNeighs;  // [sum V] reference convention: absolute fneigh index
    const float4* apars;   // [sum M]
    const float4* bLs;
    const float4* bKs;
    const float4* Ksp;
    const float4* Kpp;
    const unsigned* excl;  // exclusion bit-matrix rows
    const unsigned* excl_rot;  // rotated rows: bit k of row i = pair (i, (i+k) mod N)
    const PairCorrDev* corr;
    float4* fneigh;        // [8 * sum M] (split path only)
    float* ebond;          // [sum M] (split path only)
    double* enb;           // [sum N] per-atom non-bonded + grid energy (split path only)
    // per-system results
    int* nsteps;
    int* status;
    double* Esys;
    float* Fmax;
    float4* fire;          // per-system FIRE state (dt, damping, lastNeg, -) for the split path
    float4* stall;         // per-system learned-form fallback state (best max|F|^2, count, fallen, -)
    // preconditioner / proposal network inputs
    const int* comp;       // [sum N] connected-component index of each atom (< PC_MAXC)
    const int* ncomp;      // [nsys] number of components
    const int* elem;       // [sum N] element class 0..3 (C, N, O, H)
    // trajectory snapshots (training data): positions of every vec at steps
    // 0, snap_every, 2*snap_every, ... into snap[k * ntotV + v0 + v], k < nsnap
    float4* snap;
    int snap_every, nsnap, ntotV;
    // per-atom GridFF window cache of the persistent kernel: 16 float4 per atom,
    // row r of atom a at gridcache[r * gridcache_stride + a] (coalesced over atoms)
    float4* gridcache;
    int gridcache_stride;
    // positional restraints and bounding walls (updateAtomsMMFFf4,
    // relax_multi.cl:997-1009): per atom constr = (x, y, z, on if > 0) and
    // constrK = per-axis stiffness (negative = 0); per system bbox = 3 float4
    // (lower corner, upper corner, per-axis wall stiffness; <= 0 = no wall);
    // sysflags bit 0: the system has a restrained atom, bit 1: it has a wall
    const float4* constr;
    const float4* constrK;
    const float4* bbox;
    const int* sysflags;
    // optional per-system wall-clock record (null = off): start and end of the
    // workgroup's relaxation of system s at systime[2s], systime[2s+1]
    long long* systime;
    // learned preconditioner gains (written by the proposal network)
    const float* gnode;    // [sum M] star gains
    const float* gcomp;    // [nsys * PC_MAXC] molecule gains
    // periodic molecules: image shifts (xyz) of every system, indexed by SysInfo
    // sh0/npbc, and the image of each bonded neighbour (neighCell, per atom)
    const float4* pbcshifts;
    const int4* neighCell;
    // global-optimisation exploring (split path; null = off): per system
    // xstate = (phase: 0 relaxing / 1 exploring, steps taken in the phase,
    // minima found, -); minimum m of system s: energy minE[s * xkmax + m] and,
    // when minpos is set, positions of the system's vecs at
    // minpos[m * ntotV + v0 + v]
    int4* xstate;
    double* minE;
    float4* minpos;
};

// Parameters of the proposal network (BF16 dense layers, FP32 head).
struct MlpDev {
    const unsigned short* W1;  // [64][16]
    const unsigned short* W2;  // [64][128]
    const unsigned short* W3;  // [64][128]
    const float* b1;           // [64]
    const float* b2;
    const float* b3;
    const float* W4;           // [2][64]
    const float* b4;           // [2]
};

struct RunParams {
    FFSwitch sw;
    int bNonBonded;
    int bGridFF;
    GridDev grid;
    FireParams fire;
    float Fconv2;
    float Flimit;  // >0: clamp |f| per vec for the dynamics (relax_multi.cl:980-983); 0: off
    int conv_mode;
    int nstep_max;
    // FIRE force preconditioner: 0 off, 1 constant gains, 2 gains from the proposal network
    int pc_mode;
    float pc_gstar, pc_grigid;  // constant gains (mode 1) or gain ceilings (mode 2)
    int pc_every;               // mode 2: steps between network evaluations
    int pc_stall;               // mode 2: proposal points without a new best max|F| before a system
                                // falls back to plain FIRE (0 = never)
    MlpDev mlp;
    // chunked execution: > 0 = every system runs at most `chunk` steps per
    // launch, resuming from (and suspending to) the state in B (positions,
    // avel, fire, nsteps, status); 0 = each system runs to completion
    int chunk;
    // exploring (split path, MolWorld_sp3_multi::updateMultiExploring with the
    // Langevin drive of updateAtomsMMFFf4): a system that converges records the
    // minimum and is driven at temperature xT (K) with damping xgamma for
    // xn_explore steps, then relaxes again; a relaxation phase longer than
    // xn_relax steps (> 0) is abandoned the same way without a record
    int explore;
    float xT, xgamma;
    int xn_explore, xn_relax;
    unsigned xseed;
    int xkmax;  // minima recorded per system (further ones are counted only)
    // explicit substrate atoms instead of the GridFF (split path; nsurf = 0: none):
    // positions, REQs (R0, sqrt(E0), Q, H) and the image shifts of the substrate
    // lattice; Morse exponent surf_K = -alphaMorse
    const float4* surf_pos;
    const float4* surf_REQ;
    const float4* surf_shifts;
    int nsurf, nsurf_pbc;
    float surf_K;
};

}  // namespace fcd

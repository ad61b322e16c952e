/* libMMFFmulti_lib.so -- C-ABI of the MI355X multi-system MMFF relaxation library.
 *
 * Drop-in for the reference's libMMFFmulti_lib.so (ProkopHapala/FireCore,
 * built from cpp/libs_OCL/MMFFmulti_lib.cpp, which also compiles in
 * cpp/common/molecular/libMMFF.h and cpp/common/libUtils.h) as bound by
 * pyBall/MMFF_multi.py.  Every symbol that module binds exists here with the
 * same C signature; each declaration cites the reference function it replaces.
 *
 * Differences from the reference's behaviour are stated per function.  Common
 * to all: errors print to stderr and return a failure value (NULL / -1 / false)
 * instead of exit(0); the process-wide state is one global world (not
 * reentrant), as in the reference (MMFFmulti_lib.cpp:14).
 *
 * Additive entry points (uneven CSR batches, synthetic workloads, explicit
 * device relaxation) follow the reference ones at the end of the file.
 */
#ifndef FIRECORE_AMD_MMFF_MULTI_H
#define FIRECORE_AMD_MMFF_MULTI_H

#include <stdbool.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ======================================================= libUtils.h */
/* libUtils.h:16  print the names of all registered buffers */
void printBuffNames(void);
/* libUtils.h:23  borrowed double buffer by name ("apos","fapos","DOFs","fDOFs",
 * "vDOFs","pipos","fpipos","Es", additive: "sys_E"); NULL if unknown */
double* getBuff(const char* name);
/* libUtils.h:33  borrowed float buffer ("gpu_atoms","gpu_aforces","gpu_avel",
 * "gpu_constr","gpu_REQs","gpu_MMpars","gpu_BLs","gpu_BKs","gpu_Ksp","gpu_Kpp",
 * "gpu_lvecs","gpu_ilvecs","gpu_pbcshifts", additive: "gpu_PLQs","sys_Fmax") */
float* getfBuff(const char* name);
/* libUtils.h:43  borrowed int buffer ("neighs","gpu_neighs","gpu_neighCell",
 * "gpu_bkNeighs","ndims","selection", additive: "sys_info","sys_nsteps",
 * "sys_status","gpu_elem") */
int* getIBuff(const char* name);

/* ================================================= MMFFmulti_lib.cpp */
/* MMFFmulti_lib.cpp:23-71  (re)register the named host buffers */
void init_buffers(void);
/* MMFFmulti_lib.cpp:75-114  Build nSys replicas and upload them.
 * xyz_name: "synth:natoms=N,seed=S,sigma=X" -- template molecule from the
 *   synthetic generator, replica i perturbed with seed S+1+i.  (The reference
 *   types a molecule from an .xyz file with MMFFBuilder; that builder is not
 *   part of this library and such names return NULL.)
 * surf_name + GridFF>0: substrate "<surf_name>.xyz" (lvs line) or
 *   "synth:NaCl"; coefficients from data/<surface>/Bspline_PLQd.npy, computed
 *   and saved when missing (GridFF::tryLoad_new, GridFF.h:1787-1866).
 * nPBC must be 0 (non-periodic molecules); type-file and global-optimisation
 * arguments are accepted and ignored.  Returns the world pointer or NULL. */
void* init(int nSys, char* xyz_name, char* surf_name, char* smile_name, bool bMMFF, bool bEpairs, int* nPBC,
           int* grid_nPBC, double gridStep, char* sElementTypes, char* sAtomTypes, char* sBondTypes,
           char* sAngleTypes, double T, double gamma, int nExplore, int nRelax, double pos_kick, double vel_kick,
           int GridFF);
/* MMFFmulti_lib.cpp:116-132  Relax every system to its own Fmax < Fconv (max
 * |f| over atoms and pi-vectors) on the device; iParalel==2 uses the split
 * per-phase kernels, every other value the persistent kernel.  dt and ialg
 * are ignored (as in the reference).  outE/outF (nSys doubles each, may be
 * NULL) receive per-system energy and Fmax.  Downloads atoms/forces/vel.
 * Returns the largest step count of any system, -1 on error. */
int run(int nstepMax, double dt, double Fconv, int ialg, double* outE, double* outF, int iParalel);
/* MMFFmulti_lib.cpp:134-139  run(perframe, Ftol) */
void MDloop(int perframe, double Ftol, int iParalel, int perVF);
/* MMFFmulti_lib.cpp:141-166  FIRE parameters (cvf_min/cvf_max unused, as in the
 * FIRE path of the reference) */
void set_opt(double dt_max, double dt_min, double damp_max, double finc, double fdec, double falpha, int minLastNeg,
             double cvf_min, double cvf_max);
/* MMFFmulti_lib.cpp:168-208  GridFF probe samples along z at (pos0.x,pos0.y);
 * kinds >= 10 sample this library's B-spline grid with a probe of element
 * atyp (0=C,1=N,2=O,3=H) and charge Q; other kinds are not supported */
void sampleSurf(char* name, int n, double* rs, double* Es, double* fs, int kind, int atyp, double Q, double K,
                double RQ, double* pos0_, bool bSave);
/* MMFFmulti_lib.cpp:210-250  same at arbitrary points poss (n x 3), forces n x 3 */
void sampleSurf_vecs(char* name, int n, double* poss_, double* Es, double* fs_, int kind, int atyp, double Q,
                     double K, double RQ, double* pos0_, bool bSave);
/* MMFFmulti_lib.cpp:253-258  host system isys -> packed float mirrors */
void pack_system(int isys, bool bParams, bool bForces, bool bVel, bool blvec);
/* MMFFmulti_lib.cpp:256-258  packed float mirrors -> host system isys */
void unpack_system(int isys, bool bForces, bool bVel);
/* MMFFmulti_lib.cpp:260-262  mirrors -> device (whole batch) */
void upload_sys(int isys, bool bParams, bool bForces, bool bVel, bool blvec);
/* MMFFmulti_lib.cpp:263-265  device -> mirrors (whole batch) */
void download_sys(int isys, bool bForces, bool bVel);
/* MMFFmulti_lib.cpp:267-269 */
void upload(bool bParams, bool bForces, bool bVel, bool blvec);
/* MMFFmulti_lib.cpp:270-272 */
void download(bool bForces, bool bVel);
/* MMFFmulti_lib.cpp:275-280 (MolWorld_sp3_multi::change_lvec / add_to_lvec,
 * MolWorld_sp3_multi.h:1167-1200): bAdd = false sets the lattice of system 0
 * (the reference's current system), making it periodic with init's nPBC;
 * bAdd = true adds lvec to the lattice of every periodic system. */
void change_lvec(double* lvec, bool bAdd, bool bUpdatePi);
/* MMFFmulti_lib.cpp:290-293  not supported: no-op */
void optimizeLattice_1d(double* dlvec, int n1, int n2, int initMode, double tol);
/* MMFFmulti_lib.cpp:295-305  not supported: no-op */
void optimizeLattice_2d_multi(double* dlvec, int nstesp, int initMode, double tol);
/* MMFFmulti_lib.cpp:282-288  not supported: no-op */
void upload_pop(const char* fname);
/* MMFFmulti_lib.cpp:305-314 -> MolWorld_sp3_multi::scan_relaxed
 * (MolWorld_sp3_multi.h:1242-1296): configuration i = the template molecule
 * of init moved by rots[i] (3x3, applied transposed, Mat3::dot_to_T) and
 * poss[i], atom 0 restrained to poss[i] (stiffness 1); all nconf
 * configurations are relaxed as one batch on the device (the persistent
 * kernel's queue refills retired slots), FIRE to max|f| < Fconv, at most
 * niter_max steps, per-vec force limit Flim.  Outputs (may be NULL): Es[i]
 * force-field energy, aforces[i] (natoms x 3) force-field forces (restraint
 * removed), aposs[i] (natoms x 3) positions.  bRelax = false: no relaxation.
 * dt and omp are ignored.  The batch loaded before is restored afterwards. */
void scan(int nconf, double* poss, double* rots, double* Es, double* aforces, double* aposs, bool omp, bool bRelax,
          int niter_max, double dt, double Fconv, double Flim);
/* MMFFmulti_lib.cpp:316-329  tri-state switches (>0 on, <0 off, 0 keep);
 * bSubtractAngleNonBond follows NonBonded */
void setSwitches_multi(int CheckInvariants, int PBC, int NonBonded, int MMFF, int Angles, int PiSigma, int PiPiI,
                       int dovdW, int bSaveToDatabase);

/* ========================================================= libMMFF.h */
/* libMMFF.h:15 */
void setVerbosity(int verbosity_, int idebug_);
/* libMMFF.h:25  not supported: no-op */
void insertSMILES(char* s);
/* libMMFF.h:52  net velocity, net force and torque of system 0 */
bool checkInvariants(double maxVcog, double maxFcog, double maxTg);
/* libMMFF.h:623  not supported: no-op */
void setOptLog(int n, double* cos, double* f, double* v, double* dt, double* damp);
/* libMMFF.h:53  not supported: returns false */
bool setTrjName(const char* trj_fname_, int savePerNsteps_, int* nPBC);
/* libMMFF.h:55  all systems as concatenated xyz frames; returns #frames */
int saveXYZ(const char* fname, const char* comment, int imod);
/* libMMFF.h:112  element symbol of template atom ia */
const char* getTypeName(int ia, bool fromFF);
/* libMMFF.h (eval)  forces/energies of the batch on the device; returns E of system 0 */
double eval(void);
/* libMMFF.h:41  constraints not supported: no-op */
void addDistConstrain(int i0, int i1, double lmin, double lmax, double kmin, double kmax, double flim, double* shift,
                      bool bOldIndex);
/* libMMFF.h:46  constraints not supported: no-op */
void addAngConstrain(int i0, int i1, int i2, double ang0, double k);
/* libMMFF.h:122  shift selected atoms of system 0 */
void shift_atoms_ax(int n, int* selection, double* d);
/* libMMFF.h:124 */
void shift_atoms(int n, int* selection, int ia0, int ia1, double l);
/* libMMFF.h:123  rotate selected atoms of system 0 about (p0, ax) */
void rotate_atoms_ax(int n, int* selection, double* p0, double* ax, double phi);
/* libMMFF.h:125 */
void rotate_atoms(int n, int* selection, int ia0, int iax0, int iax1, double phi);
/* libMMFF.h:134-150  scans not supported: no-op */
void scanTranslation_ax(int n, int* selection, double* vec, int nstep, double* Es, const char* trjName,
                        bool bAddjustCaps);
void scanTranslation(int n, int* selection, int ia0, int ia1, double l, int nstep, double* Es, const char* trjName,
                     bool bAddjustCaps);
void scanRotation_ax(int n, int* selection, double* p0, double* ax, double phi, int nstep, double* Es,
                     const char* trjName);
void scanRotation(int n, int* selection, int ia0, int iax0, int iax1, double phi, int nstep, double* Es,
                  const char* trjName);
/* libMMFF.h:426  constraints not supported: zeros */
void sample_DistConstr(double lmin, double lmax, double kmin, double kmax, double flim, int n, double* xs, double* Es,
                       double* Fs);
/* libMMFF.h:439  pi-pi alignment term sampled by the device (FP32) */
void sample_evalPiAling(double k, double ang0, double r1, double r2, int n, double* angles, double* Es, double* Fs);
/* libMMFF.h:451  pi-sigma term sampled by the device (FP32) */
void sample_evalAngleCos(double k, double ang0, double r1, double r2, int n, double* angles, double* Es, double* Fs);
/* libMMFF.h:463  sigma-sigma cos(theta/2) term sampled by the device (FP32) */
void sample_evalAngleCosHalf(double k, double ang0, double r1, double r2, int n, double* angles, double* Es,
                             double* Fs);
/* libMMFF.h:475  kind 2: this library's getLJQH pair term on the device
 * (REQij mixed from REQi/REQj); other kinds not supported (zeros) */
void sampleNonBond(int n, double* rs, double* Es, double* fs, int kind, double* REQi_, double* REQj_, double K,
                   double Rdamp);

/* ============================================= additive (no reference) */
/* Uneven batch from CSR-packed host arrays: natoms/nnode per system; rows of
 * apos ((N+M) x 3), REQs/neighs (N x 4), apars..Kpp (M x 4), elem (N, may be
 * NULL) concatenated in system order.  Returns total atoms or -1. */
int init_systems(int nSys, const int* natoms, const int* nnode, const double* apos, const double* REQs,
                 const int* neighs, const double* apars, const double* bLs, const double* bKs, const double* Ksp,
                 const double* Kpp, const int* elem);
/* Synthetic workloads: kind 0 = replicas of one nmin-atom molecule, kind 1 =
 * independent molecules with N ~ U[nmin,nmax], kind 2 = molecule pairs;
 * surf_name (NULL = vacuum) places systems 3.0+-0.3 A above the substrate. */
int init_workload(int kind, int nSys, long long seed0, int nmin, int nmax, double sigma, const char* surf_name,
                  double gridStep);
/* Initial unit pi-vectors of the nnode nodes (pipos, nnode x 3) from the atom
 * positions (apos, natoms x 3), neighbour lists and the Ksp/Kpp stiffnesses:
 * MMFFsp3_loc::initPi (MMFFsp3_loc.h:882-945) without periodic images, then
 * with flip != 0 flipPis((1,1,1)) (:1167-1172).  init_systems applies it to
 * every system whose pi-vector rows are all zero.  0 on success, -1 on bad
 * input. */
int init_pi_vectors(int natoms, int nnode, const double* apos, const int* neighs, const double* Ksp,
                    const double* Kpp, int flip, double* pipos);
/* Load (or create and save) the GridFF of surf_name and attach it to the
 * current batch, as init(surf_name=...) does at initialisation (GridFF.h:
 * 1787-1866 tryLoad_new).  0 on success, -1 on error. */
int attach_surface(const char* surf_name, double gridStep);
/* The substrate as explicit atoms instead of the GridFF (MolWorld_sp3::loadSurf
 * with bGrid = false; per atom GridFF::evalMorsePBC, GridFF.h:1229-1261: Morse
 * + Coulomb pairs with every substrate atom at its images -nPBC..nPBC, null =
 * (1,1,0)), as init(surf_name, GridFF <= 0) selects.  surf_name = null removes
 * them.  Systems over substrate atoms are relaxed by the split kernels.  0 on
 * success, -1 on error. */
int attach_surface_atoms(const char* surf_name, const int* nPBC3);
/* Substrate atoms: those of surf_name read on the host (no device work), or
 * with surf_name = null the ones in use.  Positions (n x 3), REQs (n x 4),
 * lattice rows (9) and image range (3, in-use atoms only), at most maxn atoms.
 * Returns n (0: none in use), -1 if surf_name cannot be read. */
int get_surface_atoms(const char* surf_name, double* apos, double* REQs, double* lvec9, int* nPBC3, int maxn);
/* System i of init_workload (kinds 1,2) generated on the host only. */
int gen_workload_arrays(int kind, int i, long long seed0, int nmin, int nmax, double sigma, const double* pos0,
                        const double* cell, int* natoms_out, int* nnode_out, double* apos, double* REQs, int* neighs,
                        double* apars, double* bLs, double* bKs, double* Ksp, double* Kpp, int* elem);
/* Relax the batch on the device: path 0 persistent kernel, 1 split kernels;
 * restore_start restarts from the last uploaded positions; bDownload copies
 * the state back to the host mirrors.  Returns max steps, -1 on error. */
int relax(int nstepMax, double Fconv, int path, bool restore_start, bool bDownload);
/* relax() on the persistent kernel (path 0) without any device allocation.
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The relaxation's scratch (queue counters and order, activation
 * slots of the in-kernel network, GridFF window cache: 256 B per atom) lives
 * in a static device array of 64 MiB per device, held until the process ends.
 * Returns -1 with a message when the batch needs more scratch (about 250,000
 * atoms) or a path that allocates (periodic systems, systems above 256 atoms,
 * asymmetric exclusions, the chunked learned form, per-system timing).  Calls that
 * overlap on one device share its scratch and race.  Results are those of
 * relax(path 0). */
int relax_noalloc(int nstepMax, double Fconv, bool restore_start, bool bDownload);
/* Global-optimisation exploring (MolWorld_sp3_multi::updateMultiExploring,
 * MolWorld_sp3_multi.h:770-833, with the TDrive Langevin term of
 * updateAtomsMMFFf4, relax_multi.cl:1160-1178; the reference enables it with
 * init(T, gamma, nExplore, nRelax) and runs it inside run()).  Every system
 * alternates FIRE relaxation and nExplore steps of Langevin dynamics at
 * temperature T (K) and damping gamma, for nstepMax steps from its current
 * state; a relaxation longer than nRelax steps (> 0) is abandoned.  Each
 * converged relaxation records a minimum (energy; positions if bPositions),
 * up to kmax per system.  The random force is a counter-based hash of (seed,
 * system, vec, axis, step).  Returns the minima found over all systems, -1 on
 * error.  run() takes this path when init() was given nExplore > 0. */
int explore(int nstepMax, double Fconv, double T, double gamma, int nExplore, int nRelax, long long seed, int kmax,
            bool bPositions);
/* Minima of system isys from the last explore(): energies (eV) and, when
 * recorded, positions (natoms+nnode rows of x, y, z each); at most maxn.
 * Returns the number found (may exceed kmax), -1 on error. */
int get_minima(int isys, double* Es, double* poss, int maxn);
/* Minima database of the exploring mode, replaces MolecularDatabase
 * (MolecularDatabase.h:430-474, filled in MolWorld_sp3_multi.h:863-879 while
 * setSwitches_multi(..., bSaveToDatabase) is on; the reference writes new
 * members to gopt.xyz): the converged minima of explore()/run() are offered
 * to it; a structure duplicates a member when their descriptors (principal
 * axes, atom count) differ by < 5 and, optimally superposed, the root of the
 * summed squared nearest-same-type distances is < 0.1 per atom.
 * get_database_member(i<0) returns the member count; else member i's atom
 * count, with energy, system, occurrence count and positions (natoms x 3,
 * at most maxatoms rows) written where not null; -1 when out of range. */
int get_database_member(int i, double* E, int* isys, int* occurrence, double* pos, int maxatoms);
/* Offer a structure of n atoms (positions n x 3, a type label per atom): the
 * member index it duplicates, or -1 when added.  Host only. */
int database_add(int n, const double* pos, const int* types, double E);
void clear_database(void);
/* Descriptor (sqrt of the ascending inertia eigenvalues about the centroid
 * for n >= 3, then n); returns its length (at most 4). */
int database_descriptor(int n, const double* pos, double* d4);
/* The distance the database compares with 0.1 * n (see above). */
double database_distance(int n, const double* x, const int* tx, const double* y, const int* ty);
/* Members as xyz frames; returns the member count, -1 when the file cannot be written. */
int save_database_xyz(const char* fname);
/* Periodic molecules (MMFFsp3_loc with bPBC: the non-bonded sum runs over the
 * image shifts of evalPBCshifts, MolWorld_sp3.h:756-765, as
 * NBFF::evalLJQs_ng4_PBC_atom_omp, NBFF.h:575-620; bonded neighbours sit at the
 * image makeNeighCells picks, MMFFsp3_loc.h:1394-1434).  Sets the lattice rows
 * lvec9 (a, b, c) and the image range nPBC3 (null: the system's, else init's
 * nPBC) of system isys, or of all systems if isys < 0; lvec9 = null makes them
 * vacuum molecules again.  The batch is re-packed and uploaded; periodic
 * systems are relaxed by the split kernels.  0 on success, -1 on error.  No
 * counterpart in the reference's C-ABI (whose periodic cell comes from the
 * molecule's .xyz lvs line and change_lvec) beyond change_lvec, which is
 * served through it. */
int set_lattice(int isys, const double* lvec9, const int* nPBC3);
/* Device time of the last relax (HIP events), ms. */
double get_last_ms(void);
/* Queue order of the persistent kernel (longest first) by the estimated
 * cost steps(N) * N^2 with steps(N) = c0 + c1*N; feedback != 0 uses instead a
 * system's step count from the previous relaxation of the same batch when
 * there is one.  0 on success, -1 on non-finite coefficients. */
int set_cost_model(int feedback, double c0, double c1);
/* Workgroup widths of the persistent kernel that systems may use: bit c
 * enables 64*(c+1) threads (the 256-thread width is always enabled); a system
 * goes to the narrowest enabled width that holds it.  Returns the previous
 * mask, -1 if mask is outside 0..15. */
int set_class_mask(int mask);
/* Waits for all work on the device; 0, or the HIP error code of a failure no
 * earlier call has reported (message on stderr).  No counterpart in the
 * reference (its OpenCL queue reports errors per call); a test harness calls it
 * after each device test so that an asynchronous fault is charged to the test
 * that caused it. */
int device_status(void);
/* Record each system's start and end (device wall clock) in later persistent
 * relaxations (1) or not (0, default).  Returns the previous setting. */
int set_record_times(int on);
/* Start and end of every system in the last recorded relaxation, us after the
 * earliest start, into out[2*s], out[2*s+1].  0 on success, -1 if none. */
int get_system_times(double* out);
/* Persistent-kernel size classes: {block, nsys, nmax, mmax, ngroups, lds}
 * per class and each launch's device time (ms) in the last relax. */
int get_class_info(int* info6, double* ms, int maxc);
/* System indices of size class c in queue order. */
int get_class_order(int c, int* out, int maxn);
/* Forces and energies of all systems at the current positions (no move). */
int eval_forces(void);
/* Convergence criterion: 0 = max |f| < Fconv, 1 = sum f^2 < Fconv^2. */
void set_conv_mode(int mode);
/* Per-vec force clamp for the FIRE dynamics (0 = off, the CPU path's
 * behaviour; the reference GPU update kernel uses 10 eV/A,
 * relax_multi.cl:980-983).  Convergence is tested on the unclamped force. */
void set_force_limit(double Flimit);
/* FIRE force preconditioner (no counterpart in the reference, whose GPU and
 * CPU paths integrate the raw force): 0 off, 1 rigid-group projections with
 * constant gains (gstar per bonded star, grigid per molecule), 2 gains from
 * the proposal network (ceilings gstar, grigid).  0 on success, -1 invalid. */
int set_precond(int mode, double gstar, double grigid);
/* Exact forces (gpu_aforces) and preconditioned forces (view gpu_pcforces)
 * of every system at the current positions (no counterpart in the
 * reference).  0 on success. */
int eval_precond(void);
/* ---- proposal network (no counterpart in the reference) ----
 * A bonded-graph message-passing MLP (16 -> 64 -> 64 -> 64 -> 2, BF16 dense
 * layers on the matrix cores) that predicts the preconditioner gains from
 * per-atom invariants of the force, pi-orbital and substrate force.
 * Parameters: FP32, packing W1[64][16] b1 W2[64][128] b2 W3[64][128] b3
 * W4[2][64] b4[2].  All return 0 on success, -1 on error. */
int get_proposal_nparam(void);
int set_proposal_weights(const float* theta, int n);
int get_proposal_weights(float* theta, int n);
int set_proposal_every(int every);
// FIRE fallback of the learned form: after `windows` proposal points in a row
// without a new lowest max|F| a system continues with plain FIRE (zero gains,
// no more network evaluations); 0 = never.  No counterpart in the reference
// (which has no learned form).  0 on success.
int set_proposal_stall(int windows);
// Per system, 1 when the last learned relaxation fell back to plain FIRE.
int get_proposal_fallback(int* out);
/* trajectory snapshots of the next relax() (every `every` steps, at most n) */
int record_snapshots(int every, int n);
int load_snapshots(const int* idx);
/* forward pass at the current positions: out [ntotA][2], feats [ntotA][16] */
int proposal_forward(float* out, float* feats);
/* loss of one preconditioned step and its gradient w.r.t. the outputs */
int proposal_loss_grad(double eta, double dmax, float* dout, float* loss);
/* backward pass for given dloss/doutputs; grad = mean over systems */
int proposal_backward(const float* dout, float* grad);
/* Adam training on the recorded snapshots, one snapshot per system per step */
int train_proposal(int iters, double lr, double eta, double dmax, long long seed, float* loss_hist);
/* 2 = proposal network inside the persistent kernel, activations spill to global (default);
 * 1 = the same with LDS activations where the LDS budget allows; 0 = chunked relaunches */
int set_learned_inkernel(int on);
/* chunk iterations of the last learned relaxation */
int get_last_chunks(void);
/* 1 = replay the learned chunk loop from a captured HIP graph, 0 = direct launches */
int set_graphs(int on);
/* Persistent-path scheduling (no counterpart in the reference): 0 = slots
 * split between size classes by estimated time with work stealing, 1 = every
 * class at full device width, no stealing, 2 = every class at twice its
 * resident width (the extra workgroups start as other classes retire).
 * Returns the previous mode, or -1 for an unknown mode. */
int set_sched(int mode);
/* nSys, ntotA, ntotM, ntotV, maxN, maxM, grid nx, ny, nz */
void get_batch_sizes(int* out9);
/* grid pos0[3], step[3], alphaMorse; returns number of coefficients */
int get_grid_info(double* out7);
double* get_grid_coefs(void);
/* Host-only helpers (no device work): generator, topology, synthetic grid. */
int gen_molecule_arrays(int natoms, long long seed, double sigma, long long pseed, int* nnode_out, double* apos,
                        double* REQs, int* neighs, double* apars, double* bLs, double* bKs, double* Ksp, double* Kpp,
                        int* elem);
int make_back_neighs(int natoms, int nnode, const int* neighs, int* bkneighs);
/* Typed molecule of an .xyz file, built on the host as init(xyz_name) builds
 * its template (the reference's MolWorld_sp3::init with bMMFF: MMFFBuilder
 * typing, bonds, toMMFFsp3_loc, MolWorld_sp3.h:309-357, 1143-1440), with the
 * type tables fElem/fAtom/fBond (ElementTypes.dat, AtomTypes.dat,
 * BondTypes.dat).  Outputs in the MMFFsp3_loc layout, sized for maxn atoms:
 * apos 2*maxn x 3 (atoms, then one pi-vector per node), REQs maxn x 4, neighs
 * maxn x 4, apars/bLs/bKs/Ksp/Kpp maxn x 4, elem maxn; lvec9 and *periodic
 * report an lvs line.  Returns natoms, -1 on error, -2 above maxn atoms
 * (nothing written). */
int build_xyz_arrays(const char* xyz, const char* fElem, const char* fAtom, const char* fBond, int maxn,
                     int* nnode_out, double* apos, double* REQs, int* neighs, double* apars, double* bLs, double* bKs,
                     double* Ksp, double* Kpp, int* elem, double* lvec9, int* periodic);
/* Periodic molecule: image shifts in evalPBCshifts order (MolWorld_sp3.h:
 * 756-765) into shifts_out (npbc x 3) and the image of each bonded neighbour
 * (MMFFsp3_loc::makeNeighCells, MMFFsp3_loc.h:1394-1434) into neighCell_out
 * (natoms x 4, -1 for empty slots).  Returns npbc. */
int make_neigh_cells(int natoms, const double* apos, const int* neighs, const double* lvec9, const int* nPBC3,
                     double* shifts_out, int* neighCell_out);
int make_exclusions(int natoms, int nnode, const int* neighs, int subtract13, unsigned* mask_out, int* W_out,
                    int* ncorr_out);
/* Synthetic NaCl(001) slab GridFF (host sums + fit); Rdamp = Coulomb damping
 * of the node sums (0 = bare; the benchmark substrate "synth:NaCl" uses 1 A). */
int make_nacl_gridff(int nx, int ny, int nz, double dz, double Rdamp, int nthreads, double* coefs_out,
                     double* geom_out7);
/* GridFF generation (the reference's GridFF::evalBsplineRef node potentials,
 * GridFF.h:848-890: Pauli and London over the Morse images within 20 A, bare
 * Coulomb over images -25..25, r^2 >= 1e-4 A^2; then exact cubic B-spline
 * interpolation, periodic in x,y, open in z).  init()/init_workload compute the
 * node potentials on the device and fit on the host.  Substrates loaded from
 * .xyz files get the bare Coulomb channel; the synthetic benchmark substrate
 * ("synth:NaCl") the damped one, k*q/sqrt(r^2 + 1 A^2).
 *   gridff_point_values: (P, L, Q) of a substrate at npts points, host FP64
 *   gridff_nacl_nodes:   node potentials of the synthetic NaCl(001) grid with
 *                        Coulomb damping Rdamp (0 = bare) on the device
 *                        (device=1) or the host (0)
 *   gridff_fit_values:   interpolation coefficients of node values V */
int gridff_point_values(int natoms, const double* apos, const double* REQs, const double* lvec9, double alpha,
                        int npts, const double* pts, double* out);
int gridff_nacl_nodes(int nx, int ny, int nz, double dz, double Rdamp, int device, double* out);
int gridff_fit_values(int nx, int ny, int nz, const double* V, double* coefs);

#ifdef __cplusplus
}
#endif

#endif

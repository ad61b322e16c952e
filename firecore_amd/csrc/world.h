// The process-wide relaxation world behind the C-ABI (the counterpart of the
// reference's global MolWorld_sp3_multi W, cpp/libs_OCL/MMFFmulti_lib.cpp:14).
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "fc_host.h"
#include "kernels.h"

namespace fcw {

template <class T>
struct DevArr {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t cnt);
    void free_();
    void upload(const std::vector<T>& h, hipStream_t s);
    void download(std::vector<T>& h, hipStream_t s) const;
};

// HIP events of one call, destroyed on every exit path of that call.
struct EventSet {
    std::vector<hipEvent_t> ev;
    EventSet() = default;
    EventSet(const EventSet&) = delete;
    EventSet& operator=(const EventSet&) = delete;
    hipEvent_t make(unsigned flags = hipEventDefault) {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, flags) == hipSuccess) ev.push_back(e);
        return e;
    }
    ~EventSet() {
        for (hipEvent_t e : ev) hipEventDestroy(e);
    }
};

// Size class of the persistent kernel: all systems whose (N, M) fit one
// workgroup shape share one launch.
struct SizeClass {
    int block = 256;
    int cls = 0;   // index among the four widths (queue index)
    int occ = 1;   // resident workgroups per CU
    int nmax = 0, mmax = 0, wmax = 0;
    std::vector<int> order;  // system indices, decreasing cost
    double work = 0;         // sum of N^2
    int ngroups = 0;
    size_t lds = 0;
    int h2lds = 0;      // learned gains: second activation buffer in LDS (else global)
    float last_ms = 0;  // device time of this class's launch in the last relax
};

struct World {
    int verbosity = 0;
    // ---- switches (setSwitches_multi, MMFFmulti_lib.cpp:316-329)
    fcd::FFSwitch sw{1, 1, 1, 1};
    bool bNonBonded = true, bMMFF = true, bGridFF = false, bPBC = false;
    bool bCheckInvariants = false, bSaveToDatabase = false, dovdW = true;
    // ---- FIRE (MolWorld_sp3 opt after initOpt(dt_default=0.05), DynamicOpt.h:47-95)
    fcd::FireParams fp{0.05f, 0.005f, 0.1f, 1.1f, 0.5f, 0.8f, 0.0f, 3};
    int conv_mode = fcd::CONV_FMAX;
    float Flimit = 0.f;  // per-vec force clamp for the dynamics (0 = off)
    // FIRE force preconditioner (precond.h): mode 0 off, 1 constant gains, 2 learned gains
    int pc_mode = 0;
    float pc_gstar = 3.f, pc_grigid = 63.f;
    int pc_every = 8;
    int pc_stall = 0;  // learned form: proposal points without a new best max|F| before plain FIRE
    // proposal network parameters: FP32 master copy in the packing order
    // W1[64][16] b1[64] W2[64][128] b2[64] W3[64][128] b3[64] W4[2][64] b4[2]
    std::vector<float> mlp_theta;
    DevArr<unsigned short> d_mlpW;  // bf16 W1 | W2 | W3
    DevArr<float> d_mlpF;           // fp32 b1 | b2 | b3 | W4 | b4
    bool set_mlp(const float* theta, int n);
    // ---- proposal training (train.cpp)
    DevArr<float4> d_snap, d_x0, d_F0;
    int snap_every = 0, nsnap = 0;
    bool snap_on = false;
    DevArr<double> d_E0;
    DevArr<unsigned short> d_X, d_hs;
    DevArr<float> d_z, d_out, d_dout, d_gnode, d_gcomp, d_loss, d_gpart, d_grad, d_theta, d_am, d_av, d_lossmean;
    DevArr<int> d_snapidx;
    int adam_t = 0;
    bool theta_dev = false;  // d_theta holds the current parameters
    bool record_snapshots(int every, int n);
    bool load_snapshots(const std::vector<int>& idx);
    bool prop_alloc();
    bool prop_prepare();
    bool prop_forward();
    bool prop_loss_grad(float eta, float dmax);
    bool prop_backward();
    bool prop_adam(float lr);
    bool download_theta();
    bool restore_positions();
    int relax_learned(int nstep_max, double Fconv, bool restore_start);
    DevArr<int> d_count;
    bool use_graphs = true;
    bool learned_inkernel = true;  // network inside the persistent kernel (else chunked relaunches)
    bool h2lds_allowed = false;    // second activation buffer of the in-kernel network may use LDS
    DevArr<unsigned short> d_hslot;
    int last_chunks = 0;

    // ---- systems (host, double) and the template system (system 0)
    std::vector<fc::MolSys> mols;
    fc::MolSys ffl;
    fc::MolSys templ;         // the template molecule of init() (scan places copies of it)
    bool has_templ = false;
    int nSys = 0, ntotA = 0, ntotM = 0, ntotV = 0, maxN = 0, maxM = 0;
    int nvecs_uniform = 0, natoms_uniform = 0, nnode_uniform = 0;  // >0 for uniform batches
    std::vector<fcd::SysInfo> info;

    // ---- host mirrors in the reference's packed float layout
    std::vector<float4> pcforces;  // preconditioned forces of the last eval_precond
    std::vector<float4> atoms, aforces, avel, REQs, PLQs, MMpars, BLs, BKs, Ksp, Kpp, constr, lvecs, ilvecs, pbcshifts;
    // restraints of the update kernel: constrK per atom (per-axis stiffness), bboxes
    // per system (lower corner, upper corner, per-axis wall stiffness), flags per system
    std::vector<float4> constrK, bboxes;
    std::vector<int> sysflags;
    void update_sysflags();
    std::vector<int4> neighs, neighCell, bkNeighs;
    std::vector<unsigned> excl, excl_rot;
    bool ring_ok = true;
    std::vector<fcd::PairCorrDev> corr;
    std::vector<int> node_sys, elem, comp, ncomp;
    std::vector<int> nsteps, status;
    std::vector<double> Esys;
    std::vector<float> Fmax;
    // template-system views (double) exposed by getBuff
    int ndims[6] = {0, 0, 0, 0, 0, 0};
    double Es[6] = {0, 0, 0, 0, 0, 0};
    std::vector<double> DOFs, fDOFs, vDOFs;
    std::vector<int> tmpl_neighs, selection;

    // ---- GridFF
    fc::GridFFHost grid;
    std::vector<float4> grid_f4;
    // ---- explicit substrate atoms (instead of the GridFF; split kernels)
    std::vector<float4> surf_pos, surf_REQ, surf_shifts;
    fc::Surface surf_host;  // the substrate atoms in double precision
    int surf_nPBC[3] = {0, 0, 0};
    DevArr<float4> d_surf_pos, d_surf_REQ, d_surf_shifts;
    double surf_alpha = 1.5;
    // substrate atoms of s (lattice s.lvec) with images -nPBC..nPBC; replaces
    // the GridFF of the batch; s.natoms() == 0 removes them
    bool set_surface_atoms(const fc::Surface& s, const int* nPBC);

    // ---- device
    bool dev_ready = false;
    hipStream_t stream = nullptr;
    std::vector<hipStream_t> class_streams;
    DevArr<float4> d_pcforce;
    DevArr<float4> d_gridcache;  // per-atom GridFF window cache (persistent kernel)
    DevArr<float4> d_atoms, d_atoms0, d_aforce, d_avel, d_REQs, d_PLQs, d_apars, d_bLs, d_bKs, d_Ksp, d_Kpp, d_fneigh,
        d_fire, d_stall, d_grid;
    DevArr<int4> d_neighs, d_bkNeighs, d_neighCell;
    DevArr<float4> d_pbcshifts;
    int n_periodic = 0;  // periodic molecules in the batch (relaxed by the split kernels)
    int nPBC_init[3] = {0, 0, 0};  // image range given to init()
    // lattice (rows a, b, c) and image range of system isys (all if < 0); lvec
    // null makes it a vacuum molecule again.  Re-packs and uploads the batch.
    bool set_lattice(int isys, const double* lvec, const int* nPBC);
    DevArr<unsigned> d_excl, d_excl_rot;
    DevArr<fcd::PairCorrDev> d_corr;
    DevArr<fcd::SysInfo> d_info;
    DevArr<int> d_node_sys, d_nsteps, d_status, d_order, d_queue, d_comp, d_ncomp, d_elem, d_sysflags;
    DevArr<float4> d_constr, d_constrK, d_bboxes;
    DevArr<double> d_Esys, d_enb;
    DevArr<float> d_Fmax, d_ebond;
    std::vector<SizeClass> classes;
    std::vector<int> qinfo, sys_class;
    int n_unclassed = 0;  // systems too large for every class (relaxed by the split kernels)
    bool record_times = false;  // per-system start/end wall clock of the persistent kernel
    DevArr<long long> d_systime;
    DevArr<int> d_only;
    bool hb_batch = false;
    bool autotune = true;
    // cost estimate of a system for the queue order (longest first):
    //   feedback: the step count of its previous relaxation when there is one,
    //   else the model steps = cost_c0 + cost_c1 * N; model only: never uses
    //   step counts of earlier relaxations of the same batch
    bool cost_feedback = true;
    int class_mask = 0xF;  // enabled workgroup widths: bit c = 64*(c+1) threads (256 always enabled)
    double cost_c0 = 1000.0, cost_c1 = 20.0;
    // 0: device slots split between classes by estimated time, idle workgroups steal
    //    from smaller classes; 1: every class launched at full device width, no stealing
    int sched_mode = 1;
    double last_ms = 0;

    // ---- buffer registry (libUtils.h:16-66)
    std::map<std::string, double*> buffers;
    std::map<std::string, float*> fbuffers;
    std::map<std::string, int*> ibuffers;

    void clear();
    bool set_systems(std::vector<fc::MolSys>&& ms);
    void build_packed();
    void rebuild_exclusions();
    void pack_system(int isys, bool bParams, bool bForces, bool bVel, bool blvec);
    void unpack_system(int isys, bool bForces, bool bVel);
    bool ensure_device();
    bool upload_all(bool bParams, bool bForces, bool bVel);
    bool download_all(bool bForces, bool bVel);
    bool set_grid(const fc::GridFFHost& g);
    void register_buffers();
    fcd::BatchDev batch_dev();
    fcd::RunParams run_params(double Fconv, int nstep_max);
    // launch plan of the persistent kernel; upload = false leaves the queue
    // order in h_order only (the allocation-free path copies it itself)
    void plan_classes(bool upload = true);
    std::vector<int> h_order;
    // device entry points
    int relax_persistent(int nstep_max, double Fconv, bool restore_start);
    int relax_split(int nstep_max, double Fconv, bool restore_start);
    // relax_persistent with every scratch buffer in the library's static device
    // array (no allocation); -1 with a message when the batch needs more scratch
    // than FC_NOALLOC_SCRATCH_BYTES or a path that allocates
    int relax_noalloc(int nstep_max, double Fconv, bool restore_start);
    void* noalloc_ptr = nullptr;  // this device's scratch array (looked up at set-up)
    std::vector<hipEvent_t> noalloc_events;  // fork/join events of the class launches, made at set-up
    bool relax_unclassed(int nstep_max, double Fconv);
    // screening caller (MolWorld_sp3_multi::scan_relaxed): every configuration
    // a system of one batch, relaxed together; the previous batch is restored
    bool scan(int nconf, const double* poss, const double* rots, double* Es, double* aforces, double* aposs,
              bool bRelax, int niter_max, double Fconv, double Flim);
    // global-optimisation exploring (split kernels): minima per system
    int explore(int nstep_max, double Fconv, double T, double gamma, int n_explore, int n_relax, unsigned seed,
                int kmax, bool keep_positions);
    DevArr<int4> d_xstate;
    DevArr<double> d_minE;
    DevArr<float4> d_minpos;
    std::vector<double> minE;   // [nSys * xkmax]
    std::vector<float4> minpos;  // [xkmax * ntotV]
    std::vector<int> nmin;       // minima found per system (may exceed xkmax)
    int xkmax = 0;
    bool minpos_kept = false;
    // minima database (filled by explore() while bSaveToDatabase is on)
    fc::MinimaDB db;
    void feed_database();
    // exploring parameters of init() (T, gamma, nExplore, nRelax); run() explores when nExplore > 0
    double go_T = -1, go_gamma = -1;
    int go_nExplore = 0, go_nRelax = 0;
    bool eval_forces();
    bool eval_precond();
    bool sync();
};

World& world();
bool hip_ok(hipError_t e, const char* what);

}  // namespace fcw

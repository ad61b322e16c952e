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

    // ---- systems (host, double) and the template system (system 0)
    std::vector<fc::MolSys> mols;
    fc::MolSys ffl;
    int nSys = 0, ntotA = 0, ntotM = 0, ntotV = 0, maxN = 0, maxM = 0;
    int nvecs_uniform = 0, natoms_uniform = 0, nnode_uniform = 0;  // >0 for uniform batches
    std::vector<fcd::SysInfo> info;

    // ---- host mirrors in the reference's packed float layout
    std::vector<float4> atoms, aforces, avel, REQs, PLQs, MMpars, BLs, BKs, Ksp, Kpp, constr, lvecs, ilvecs, pbcshifts;
    std::vector<int4> neighs, neighCell, bkNeighs;
    std::vector<unsigned> excl, excl_rot;
    bool ring_ok = true;
    std::vector<fcd::PairCorrDev> corr;
    std::vector<int> node_sys, elem;
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

    // ---- device
    bool dev_ready = false;
    hipStream_t stream = nullptr;
    std::vector<hipStream_t> class_streams;
    DevArr<float4> d_atoms, d_atoms0, d_aforce, d_avel, d_REQs, d_PLQs, d_apars, d_bLs, d_bKs, d_Ksp, d_Kpp, d_fneigh,
        d_fire, d_grid;
    DevArr<int4> d_neighs, d_bkNeighs;
    DevArr<unsigned> d_excl, d_excl_rot;
    DevArr<fcd::PairCorrDev> d_corr;
    DevArr<fcd::SysInfo> d_info;
    DevArr<int> d_node_sys, d_nsteps, d_status, d_order, d_queue;
    DevArr<double> d_Esys, d_enb;
    DevArr<float> d_Fmax, d_ebond;
    std::vector<SizeClass> classes;
    std::vector<int> qinfo, sys_class;
    bool hb_batch = false;
    bool autotune = true;
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
    void plan_classes();
    // device entry points
    int relax_persistent(int nstep_max, double Fconv, bool restore_start);
    int relax_split(int nstep_max, double Fconv, bool restore_start);
    bool eval_forces();
    bool sync();
};

World& world();
bool hip_ok(hipError_t e, const char* what);

}  // namespace fcw

// The C-ABI of libMMFFmulti_lib.so: every symbol that pyBall/MMFF_multi.py
// binds (cpp/libs_OCL/MMFFmulti_lib.cpp, cpp/common/molecular/libMMFF.h,
// cpp/common/libUtils.h), with the same C signatures, plus additive entry
// points for uneven (CSR-packed) batches.  Declarations: include/mmff_multi.h.
//
// Error convention: the reference prints and calls exit(0); this library
// prints to stderr and returns a failure value (NULL, -1 or false) instead.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <string>
#include <sys/stat.h>
#include <thread>

#include "world.h"

extern "C" hipError_t fc_sample_angles(int kind, float k, float ang0, float r1, float r2, int n, const float* ang,
                                       float* E, float* F);
extern "C" hipError_t fc_sample_grid(const fcd::GridDev* g, int n, const float4* pos, float4 plq, float4* out);
extern "C" hipError_t fc_sample_pair(int n, const float* r, float4 REQij, float2* out);
extern "C" hipError_t fc_gridff_nodes_device(const fc::Surface* s, const fc::GridFFHost* g, double* V);

using fcw::World;
using fcw::world;

namespace {

void unsupported(const char* name) {
    std::fprintf(stderr, "[libMMFFmulti_lib] %s: not part of the multi-system relaxation path; no-op\n", name);
}

bool file_exists(const std::string& f) {
    struct stat st;
    return stat(f.c_str(), &st) == 0;
}
void mkdirs(const std::string& d) {
    std::string acc;
    for (size_t i = 0; i < d.size(); i++) {
        acc += d[i];
        if (d[i] == '/' || i + 1 == d.size()) mkdir(acc.c_str(), 0755);
    }
}
std::string basename_noext(const std::string& p) {
    std::string b = p;
    size_t k = b.find_last_of('/');
    if (k != std::string::npos) b = b.substr(k + 1);
    if (b.size() > 4 && b.substr(b.size() - 4) == ".xyz") b = b.substr(0, b.size() - 4);
    return b;
}

// Surface + GridFF set-up (MolWorld_sp3::loadSurf, MolWorld_sp3.h:1058-1084, and
// MolWorld_sp3_multi::initGridFF, MolWorld_sp3_multi.h:2501-2600): the B-spline
// coefficients are read from data/<surface>/Bspline_PLQd.npy when present and
// computed and saved there otherwise (GridFF::tryLoad_new, GridFF.h:1787-1866).
bool setup_surface(const char* surf_name, double gridStep, fc::GridFFHost& g) {
    std::string sn = surf_name;
    fc::Surface S;
    std::string base;
    if (sn.rfind("synth:NaCl", 0) == 0) {
        fc::make_nacl_slab(5.64, 2, 0.9, S);
        S.Rdamp = fc::NACL_SYNTH_RDAMP;
        base = "NaCl_synth_Rdamp1";
    } else {
        std::string fx = sn;
        if (fx.size() < 4 || fx.substr(fx.size() - 4) != ".xyz") fx += ".xyz";
        if (!fc::load_surface_xyz(fx, S)) {
            std::fprintf(stderr, "[libMMFFmulti_lib] cannot read surface %s\n", fx.c_str());
            return false;
        }
        base = basename_noext(sn);
    }
    double ztop = -1e300;
    for (int i = 0; i < S.natoms(); i++) ztop = std::max(ztop, S.apos[i * 3 + 2]);
    const double ax = S.lvec[0], by = S.lvec[4], cz = S.lvec[8];
    const std::string dir = "data/" + base;
    const std::string fn = dir + "/Bspline_PLQd.npy";
    std::vector<int> shape;
    std::vector<double> data;
    if (file_exists(fn) && fc::npy_load_f64(fn, shape, data) && shape.size() == 4 && shape[3] == 3) {
        g.nx = shape[0];
        g.ny = shape[1];
        g.nz = shape[2];
        g.step[0] = ax / g.nx;
        g.step[1] = by / g.ny;
        g.step[2] = cz / g.nz;
        g.pos0[0] = -0.5 * ax;
        g.pos0[1] = -0.5 * by;
        g.pos0[2] = ztop;
        g.alphaMorse = 1.5;
        g.coefs = std::move(data);
        return true;
    }
    const int nx = (int)std::lround(ax / gridStep), ny = (int)std::lround(by / gridStep),
              nz = (int)std::lround(cz / gridStep);
    // node potentials on the device (FP64), exact B-spline fit on the host
    World& W = world();
    if (!W.ensure_device()) return false;
    fc::gridff_geometry(S, nx, ny, nz, cz / nz, ztop, 1.5, g);
    std::vector<double> V((size_t)nx * ny * nz * 3);
    if (!fcw::hip_ok(fc_gridff_nodes_device(&S, &g, V.data()), "GridFF node potentials")) return false;
    fc::gridff_fit(g, V, 16);
    mkdirs(dir);
    std::vector<int> sh = {g.nx, g.ny, g.nz, 3};
    if (!fc::npy_save_f64(fn, sh, g.coefs.data()))
        std::fprintf(stderr, "[libMMFFmulti_lib] warning: cannot write %s\n", fn.c_str());
    return true;
}

// Substrate atoms of surf_name ("synth:NaCl" or an .xyz with an lvs line),
// MolWorld_sp3::loadSurf with bGrid = false (MolWorld_sp3.h:1058-1084).
bool load_surface_atoms(const char* surf_name, fc::Surface& S) {
    std::string sn = surf_name;
    if (sn.rfind("synth:NaCl", 0) == 0) {
        fc::make_nacl_slab(5.64, 2, 0.9, S);
        return true;
    }
    std::string fx = sn;
    if (fx.size() < 4 || fx.substr(fx.size() - 4) != ".xyz") fx += ".xyz";
    if (!fc::load_surface_xyz(fx, S)) {
        std::fprintf(stderr, "[libMMFFmulti_lib] cannot read surface %s\n", fx.c_str());
        return false;
    }
    return true;
}

bool finish_init(World& W) {
    if (!W.ensure_device()) return false;
    if (!W.upload_all(true, true, true)) return false;
    return true;
}

}  // namespace

extern "C" {

// =============================================================== libUtils.h
void printBuffNames() {
    World& W = world();
    std::printf("buffers:  ");
    for (auto& kv : W.buffers) std::printf("%s ", kv.first.c_str());
    std::printf("\nfbuffers: ");
    for (auto& kv : W.fbuffers) std::printf("%s ", kv.first.c_str());
    std::printf("\nibuffers: ");
    for (auto& kv : W.ibuffers) std::printf("%s ", kv.first.c_str());
    std::printf("\n");
}
double* getBuff(const char* name) {
    World& W = world();
    auto it = W.buffers.find(name);
    if (it == W.buffers.end()) {
        std::printf("ERROR: getBuff(%s) NOT FOUND\n", name);
        printBuffNames();
        return nullptr;
    }
    return it->second;
}
float* getfBuff(const char* name) {
    World& W = world();
    auto it = W.fbuffers.find(name);
    if (it == W.fbuffers.end()) {
        std::printf("ERROR: getfBuff(%s) NOT FOUND\n", name);
        printBuffNames();
        return nullptr;
    }
    return it->second;
}
int* getIBuff(const char* name) {
    World& W = world();
    auto it = W.ibuffers.find(name);
    if (it == W.ibuffers.end()) {
        std::printf("ERROR: getIBuff(%s) NOT FOUND\n", name);
        printBuffNames();
        return nullptr;
    }
    return it->second;
}

// ========================================================= MMFFmulti_lib.cpp
void init_buffers() { world().register_buffers(); }

// xyz_name: a typed molecule file (".xyz" appended when missing), built as the
// reference's MolWorld_sp3::init does with bMMFF (MolWorld_sp3.h:309-357;
// typing tables from common_resources/ in the working directory as the
// reference loads them, else from the given sElementTypes/sAtomTypes/
// sBondTypes paths): nSys identical replicas, periodic with init's nPBC when
// the file has an lvs line.  "synth:natoms=<N>,seed=<s>,sigma=<A>" builds the
// template with the synthetic generator instead, replica i perturbed with
// seed s+1+i.  bEpairs (explicit electron pairs) is not supported.
void* init(int nSys, char* xyz_name, char* surf_name, char* smile_name, bool bMMFF, bool bEpairs, int* nPBC,
           int* grid_nPBC, double gridStep, char* sElementTypes, char* sAtomTypes, char* sBondTypes,
           char* sAngleTypes, double T, double gamma, int nExplore, int nRelax, double pos_kick, double vel_kick,
           int GridFF) {
    (void)smile_name; (void)sAngleTypes; (void)pos_kick; (void)vel_kick;
    World& W = world();
    // global-optimisation parameters (W.go, MMFFmulti_lib.cpp:91-96): run() explores when nExplore > 0
    W.go_T = T;
    W.go_gamma = gamma;
    W.go_nExplore = nExplore;
    W.go_nRelax = nRelax;
    for (int k = 0; k < 3; k++) W.nPBC_init[k] = nPBC ? std::max(nPBC[k], 0) : 0;
    if (!xyz_name || nSys < 1) {
        std::fprintf(stderr, "[libMMFFmulti_lib] init: needs xyz_name and nSys >= 1\n");
        return nullptr;
    }
    fc::MolSys tmpl;
    std::vector<fc::MolSys> ms;
    if (std::strncmp(xyz_name, "synth:", 6) != 0) {
        if (bEpairs) std::fprintf(stderr, "[libMMFFmulti_lib] init: explicit electron pairs are not supported; building without them\n");
        auto pick = [](const char* given, const char* deflt) {
            if (file_exists(deflt) || !given) return std::string(deflt);
            return std::string(given);
        };
        fc::TypeTables tables;
        std::string err;
        if (!tables.load(pick(sElementTypes, "common_resources/ElementTypes.dat"),
                         pick(sAtomTypes, "common_resources/AtomTypes.dat"),
                         pick(sBondTypes, "common_resources/BondTypes.dat"), err)) {
            std::fprintf(stderr, "[libMMFFmulti_lib] init: %s\n", err.c_str());
            return nullptr;
        }
        std::string fn = xyz_name;
        if (fn.size() < 4 || fn.substr(fn.size() - 4) != ".xyz") fn += ".xyz";
        bool periodic = false;
        double lvec[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (!fc::build_from_xyz(fn, tables, periodic, lvec, tmpl, err)) {
            std::fprintf(stderr, "[libMMFFmulti_lib] init(%s): %s\n", fn.c_str(), err.c_str());
            return nullptr;
        }
        if (periodic) {  // MolWorld_sp3::makeMMFFs PBC set-up with init's nPBC
            tmpl.pbc = true;
            std::memcpy(tmpl.lvec, lvec, sizeof(lvec));
            for (int k = 0; k < 3; k++) tmpl.nPBC[k] = W.nPBC_init[k];
        }
        ms.assign(nSys, tmpl);
    } else {
        int natoms = 24;
        long long seed = 0;
        double sigma = 0.1;
        const std::string spec(xyz_name + 6);
        size_t p = 0;
        while (p < spec.size()) {
            const size_t q = spec.find(',', p);
            const std::string kv = spec.substr(p, q == std::string::npos ? std::string::npos : q - p);
            const size_t e = kv.find('=');
            if (e != std::string::npos) {
                const std::string k = kv.substr(0, e), v = kv.substr(e + 1);
                if (k == "natoms") natoms = std::atoi(v.c_str());
                else if (k == "seed") seed = std::atoll(v.c_str());
                else if (k == "sigma") sigma = std::atof(v.c_str());
            }
            if (q == std::string::npos) break;
            p = q + 1;
        }
        if (!fc::gen_molecule(natoms, (uint64_t)seed, tmpl)) {
            std::fprintf(stderr, "[libMMFFmulti_lib] init: cannot generate a %d-atom molecule\n", natoms);
            return nullptr;
        }
        ms.assign(nSys, tmpl);
        for (int i = 0; i < nSys; i++) fc::perturb_positions(ms[i], (uint64_t)(seed + 1 + i), sigma);
    }
    W.templ = tmpl;
    W.has_templ = true;
    W.bMMFF = bMMFF;
    W.bGridFF = false;
    if (!W.set_systems(std::move(ms))) return nullptr;
    if (surf_name && GridFF > 0) {
        fc::GridFFHost g;
        if (!setup_surface(surf_name, gridStep > 0 ? gridStep : 0.1, g)) return nullptr;
        W.set_grid(g);
    }
    if (!finish_init(W)) return nullptr;
    if (surf_name && GridFF <= 0) {  // substrate as explicit atoms (bSurfAtoms without bGridFF)
        fc::Surface S;
        const int np0[3] = {1, 1, 0};
        if (!load_surface_atoms(surf_name, S) || !W.set_surface_atoms(S, grid_nPBC ? grid_nPBC : np0)) return nullptr;
    }
    return &W;
}

int run(int nstepMax, double dt, double Fconv, int ialg, double* outE, double* outF, int iParalel) {
    (void)dt;
    (void)ialg;
    World& W = world();
    if (!W.dev_ready || W.nSys == 0) {
        std::fprintf(stderr, "[libMMFFmulti_lib] run: not initialised\n");
        return -1;
    }
    int n;
    if (W.go_nExplore > 0) {  // exploring requested at init: relax-explore cycles on the split kernels
        n = W.explore(nstepMax, Fconv, W.go_T, W.go_gamma, W.go_nExplore, W.go_nRelax, 0x5EEDu, 16, true) < 0
                ? -1
                : nstepMax;
    } else {
        n = (iParalel == 2) ? W.relax_split(nstepMax, Fconv, false) : W.relax_persistent(nstepMax, Fconv, false);
    }
    if (n < 0) return n;
    // state stays on the device for the next call: refresh the start copy
    hipMemcpyAsync(W.d_atoms0.p, W.d_atoms.p, sizeof(float4) * W.ntotV, hipMemcpyDeviceToDevice, W.stream);
    W.download_all(true, true);
    if (outE)
        for (int s = 0; s < W.nSys; s++) outE[s] = W.Esys[s];
    if (outF)
        for (int s = 0; s < W.nSys; s++) outF[s] = W.Fmax[s];
    return n;
}

void MDloop(int perframe, double Ftol, int iParalel, int perVF) {
    (void)perVF;
    run(perframe, -1, Ftol > 0 ? Ftol : 1e-3, 0, nullptr, nullptr, iParalel);
}

void set_opt(double dt_max, double dt_min, double damp_max, double finc, double fdec, double falpha, int minLastNeg,
             double cvf_min, double cvf_max) {
    (void)cvf_min;
    (void)cvf_max;
    World& W = world();
    W.fp.dt_max = (float)dt_max;
    W.fp.dt_min = (float)dt_min;
    W.fp.damp_max = (float)damp_max;
    W.fp.finc = (float)finc;
    W.fp.fdec = (float)fdec;
    W.fp.falpha = (float)falpha;
    W.fp.minLastNeg = minLastNeg;
}

void sampleSurf(char* name, int n, double* rs, double* Es, double* fs, int kind, int atyp, double Q, double K,
                double RQ, double* pos0_, bool bSave);
void sampleSurf_vecs(char* name, int n, double* poss_, double* Es, double* fs_, int kind, int atyp, double Q,
                     double K, double RQ, double* pos0_, bool bSave);

void pack_system(int isys, bool bParams, bool bForces, bool bVel, bool blvec) {
    World& W = world();
    if (isys < 0 || isys >= W.nSys) return;
    W.pack_system(isys, bParams, bForces, bVel, blvec);
}
void unpack_system(int isys, bool bForces, bool bVel) {
    World& W = world();
    if (isys < 0 || isys >= W.nSys) return;
    W.unpack_system(isys, bForces, bVel);
}
void upload_sys(int isys, bool bParams, bool bForces, bool bVel, bool blvec) {
    (void)isys;
    (void)blvec;
    World& W = world();
    W.upload_all(bParams, bForces, bVel);
}
void download_sys(int isys, bool bForces, bool bVel) {
    (void)isys;
    world().download_all(bForces, bVel);
}
void upload(bool bParams, bool bForces, bool bVel, bool blvec) {
    (void)blvec;
    World& W = world();
    if (bParams) W.rebuild_exclusions();
    W.upload_all(bParams, bForces, bVel);
}
void download(bool bForces, bool bVel) { world().download_all(bForces, bVel); }

// MolWorld_sp3_multi::change_lvec / add_to_lvec (MolWorld_sp3_multi.h:1167-1200):
// bAdd = false sets the lattice of the current system (system 0), bAdd = true
// adds dlvec to the lattice of every periodic system; the image shifts follow.
// Systems become periodic with the image range of init's nPBC.
void change_lvec(double* lvec, bool bAdd, bool bUpdatePi) {
    (void)bUpdatePi;
    World& W = world();
    if (!lvec || W.nSys == 0) return;
    if (!bAdd) {
        if (!W.set_lattice(0, lvec, W.mols[0].pbc ? W.mols[0].nPBC : W.nPBC_init))
            std::fprintf(stderr, "[libMMFFmulti_lib] change_lvec failed\n");
        return;
    }
    for (int s = 0; s < W.nSys; s++) {
        if (!W.mols[s].pbc) continue;
        double L[9];
        for (int k = 0; k < 9; k++) L[k] = W.mols[s].lvec[k] + lvec[k];
        if (!W.set_lattice(s, L, W.mols[s].nPBC)) std::fprintf(stderr, "[libMMFFmulti_lib] change_lvec failed\n");
    }
}
void optimizeLattice_1d(double* dlvec, int n1, int n2, int initMode, double tol) {
    (void)dlvec; (void)n1; (void)n2; (void)initMode; (void)tol;
    unsupported("optimizeLattice_1d");
}
void optimizeLattice_2d_multi(double* dlvec, int nstesp, int initMode, double tol) {
    (void)dlvec; (void)nstesp; (void)initMode; (void)tol;
    unsupported("optimizeLattice_2d_multi");
}
void upload_pop(const char* fname) {
    (void)fname;
    unsupported("upload_pop");
}
void scan(int nconf, double* poss, double* rots, double* Es, double* aforces, double* aposs, bool omp, bool bRelax,
          int niter_max, double dt, double Fconv, double Flim) {
    (void)omp;
    (void)dt;
    if (!world().scan(nconf, poss, rots, Es, aforces, aposs, bRelax, niter_max, Fconv, Flim))
        std::fprintf(stderr, "[libMMFFmulti_lib] scan failed (no batch initialised, or a device error)\n");
}
void setSwitches_multi(int CheckInvariants, int PBC, int NonBonded, int MMFF, int Angles, int PiSigma, int PiPiI,
                       int dovdW, int bSaveToDatabase) {
    World& W = world();
    auto setb = [](bool& b, int i) { if (i > 0) b = true; else if (i < 0) b = false; };
    auto seti = [](int& b, int i) { if (i > 0) b = 1; else if (i < 0) b = 0; };
    setb(W.bCheckInvariants, CheckInvariants);
    setb(W.bPBC, PBC);
    setb(W.bNonBonded, NonBonded);
    setb(W.bMMFF, MMFF);
    seti(W.sw.doAngles, Angles);
    seti(W.sw.doPiSigma, PiSigma);
    seti(W.sw.doPiPiI, PiPiI);
    setb(W.dovdW, dovdW);
    setb(W.bSaveToDatabase, bSaveToDatabase);
    if (W.nSys > 0) {
        // bSubtractAngleNonBond follows bNonBonded (MMFFmulti_lib.cpp:328)
        W.rebuild_exclusions();
        if (W.dev_ready) {
            W.d_excl.upload(W.excl, W.stream);
            W.d_excl_rot.upload(W.excl_rot, W.stream);
            W.d_corr.upload(W.corr, W.stream);
            W.d_info.upload(W.info, W.stream);
            W.sync();
        }
    }
}

// ================================================================ libMMFF.h
void setVerbosity(int verbosity_, int idebug_) {
    (void)idebug_;
    world().verbosity = verbosity_;
    setbuf(stdout, NULL);
}
void insertSMILES(char* s) {
    (void)s;
    unsupported("insertSMILES");
}
bool checkInvariants(double maxVcog, double maxFcog, double maxTg) {
    // net velocity, net force and torque of the template system's atoms
    World& W = world();
    const fc::MolSys& m = W.ffl;
    double vc[3] = {0, 0, 0}, fc_[3] = {0, 0, 0}, tq[3] = {0, 0, 0};
    for (int i = 0; i < m.natoms; i++) {
        const double* p = &W.DOFs[i * 3];
        const double* f = &W.fDOFs[i * 3];
        const double* v = &W.vDOFs[i * 3];
        for (int k = 0; k < 3; k++) {
            vc[k] += v[k];
            fc_[k] += f[k];
        }
        tq[0] += p[1] * f[2] - p[2] * f[1];
        tq[1] += p[2] * f[0] - p[0] * f[2];
        tq[2] += p[0] * f[1] - p[1] * f[0];
    }
    auto n = [](const double* a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); };
    return (n(vc) <= maxVcog) && (n(fc_) <= maxFcog) && (n(tq) <= maxTg);
}
void setOptLog(int n, double* cos, double* f, double* v, double* dt, double* damp) {
    (void)n; (void)cos; (void)f; (void)v; (void)dt; (void)damp;
    unsupported("setOptLog");
}
bool setTrjName(const char* trj_fname_, int savePerNsteps_, int* nPBC) {
    (void)trj_fname_; (void)savePerNsteps_; (void)nPBC;
    unsupported("setTrjName");
    return false;
}
int saveXYZ(const char* fname, const char* comment, int imod) {
    (void)imod;
    World& W = world();
    FILE* f = std::fopen(fname, "w");
    if (!f) return -1;
    static const char* el[4] = {"C", "N", "O", "H"};
    for (int s = 0; s < W.nSys; s++) {
        const fcd::SysInfo& I = W.info[s];
        std::fprintf(f, "%d\n%s sys=%d\n", I.natoms, comment ? comment : "", s);
        for (int i = 0; i < I.natoms; i++) {
            const float4 a = W.atoms[I.v0 + i];
            std::fprintf(f, "%s %.6f %.6f %.6f\n", el[W.elem[I.a0 + i] & 3], a.x, a.y, a.z);
        }
    }
    std::fclose(f);
    return W.nSys;
}
const char* getTypeName(int ia, bool fromFF) {
    (void)fromFF;
    static const char* el[4] = {"C", "N", "O", "H"};
    World& W = world();
    if (ia < 0 || ia >= W.ffl.natoms || W.ffl.elem.empty()) return "?";
    return el[W.ffl.elem[ia] & 3];
}
double eval() {
    World& W = world();
    if (!W.dev_ready || W.nSys == 0) return NAN;
    if (!W.eval_forces()) return NAN;
    W.download_all(true, false);
    return W.Esys[0];
}
void addDistConstrain(int i0, int i1, double lmin, double lmax, double kmin, double kmax, double flim, double* shift,
                      bool bOldIndex) {
    (void)i0; (void)i1; (void)lmin; (void)lmax; (void)kmin; (void)kmax; (void)flim; (void)shift; (void)bOldIndex;
    unsupported("addDistConstrain");
}
void addAngConstrain(int i0, int i1, int i2, double ang0, double k) {
    (void)i0; (void)i1; (void)i2; (void)ang0; (void)k;
    unsupported("addAngConstrain");
}
static void apply_to_template(int n, const int* sel, void (*fn)(double*, const double*), const double* arg) {
    World& W = world();
    if (W.nSys == 0) return;
    fc::MolSys& m = W.mols[0];
    if (!sel || n <= 0) {
        for (int i = 0; i < m.natoms; i++) fn(&m.apos[i * 3], arg);
    } else {
        for (int k = 0; k < n; k++)
            if (sel[k] >= 0 && sel[k] < m.natoms) fn(&m.apos[sel[k] * 3], arg);
    }
    W.pack_system(0, false, false, false, false);
    W.ffl.apos = m.apos;
    W.DOFs = m.apos;
    W.register_buffers();
    if (W.dev_ready) W.upload_all(false, false, false);
}
void shift_atoms_ax(int n, int* selection, double* d) {
    apply_to_template(n, selection, [](double* p, const double* a) {
        p[0] += a[0];
        p[1] += a[1];
        p[2] += a[2];
    }, d);
}
void shift_atoms(int n, int* selection, int ia0, int ia1, double l) {
    World& W = world();
    if (W.nSys == 0) return;
    const fc::MolSys& m = W.mols[0];
    if (ia0 < 0 || ia1 < 0 || ia0 >= m.natoms || ia1 >= m.natoms) return;
    double d[3];
    double nn = 0;
    for (int k = 0; k < 3; k++) {
        d[k] = m.apos[ia1 * 3 + k] - m.apos[ia0 * 3 + k];
        nn += d[k] * d[k];
    }
    nn = std::sqrt(nn);
    for (int k = 0; k < 3; k++) d[k] *= l / nn;
    shift_atoms_ax(n, selection, d);
}
void rotate_atoms_ax(int n, int* selection, double* p0, double* ax, double phi) {
    double arg[7] = {p0 ? p0[0] : 0, p0 ? p0[1] : 0, p0 ? p0[2] : 0, ax ? ax[0] : 0, ax ? ax[1] : 0,
                     ax ? ax[2] : 1, phi};
    apply_to_template(n, selection, [](double* p, const double* a) {
        double u[3] = {a[3], a[4], a[5]};
        double nu = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        for (double& x : u) x /= nu;
        double v[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
        double c = std::cos(a[6]), s = std::sin(a[6]);
        double uv = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
        double cx[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        for (int k = 0; k < 3; k++) p[k] = a[k] + v[k] * c + cx[k] * s + u[k] * uv * (1 - c);
    }, arg);
}
void rotate_atoms(int n, int* selection, int ia0, int iax0, int iax1, double phi) {
    World& W = world();
    if (W.nSys == 0) return;
    const fc::MolSys& m = W.mols[0];
    if (ia0 < 0 || iax0 < 0 || iax1 < 0 || ia0 >= m.natoms || iax0 >= m.natoms || iax1 >= m.natoms) return;
    double p0[3], ax[3];
    for (int k = 0; k < 3; k++) {
        p0[k] = m.apos[ia0 * 3 + k];
        ax[k] = m.apos[iax1 * 3 + k] - m.apos[iax0 * 3 + k];
    }
    rotate_atoms_ax(n, selection, p0, ax, phi);
}
void scanTranslation_ax(int n, int* selection, double* vec, int nstep, double* Es, const char* trjName,
                        bool bAddjustCaps) {
    (void)n; (void)selection; (void)vec; (void)nstep; (void)Es; (void)trjName; (void)bAddjustCaps;
    unsupported("scanTranslation_ax");
}
void scanTranslation(int n, int* selection, int ia0, int ia1, double l, int nstep, double* Es, const char* trjName,
                     bool bAddjustCaps) {
    (void)n; (void)selection; (void)ia0; (void)ia1; (void)l; (void)nstep; (void)Es; (void)trjName;
    (void)bAddjustCaps;
    unsupported("scanTranslation");
}
void scanRotation_ax(int n, int* selection, double* p0, double* ax, double phi, int nstep, double* Es,
                     const char* trjName) {
    (void)n; (void)selection; (void)p0; (void)ax; (void)phi; (void)nstep; (void)Es; (void)trjName;
    unsupported("scanRotation_ax");
}
void scanRotation(int n, int* selection, int ia0, int iax0, int iax1, double phi, int nstep, double* Es,
                  const char* trjName) {
    (void)n; (void)selection; (void)ia0; (void)iax0; (void)iax1; (void)phi; (void)nstep; (void)Es; (void)trjName;
    unsupported("scanRotation");
}

// ---- term samplers: evaluated by the device's FP32 terms (libMMFF.h:426-497)
static void sample_angles_dev(int kind, double k, double ang0, double r1, double r2, int n, const double* angles,
                              double* Es, double* Fs) {
    World& W = world();
    if (n <= 0 || !W.ensure_device()) return;
    std::vector<float> a(n), e(n), f(n);
    for (int i = 0; i < n; i++) a[i] = (float)angles[i];
    float *da, *de, *df;
    hipMalloc(&da, n * 4);
    hipMalloc(&de, n * 4);
    hipMalloc(&df, n * 4);
    hipMemcpy(da, a.data(), n * 4, hipMemcpyHostToDevice);
    fcw::hip_ok(fc_sample_angles(kind, (float)k, (float)ang0, (float)r1, (float)r2, n, da, de, df), "sample_angles");
    hipMemcpy(e.data(), de, n * 4, hipMemcpyDeviceToHost);
    hipMemcpy(f.data(), df, n * 4, hipMemcpyDeviceToHost);
    hipFree(da);
    hipFree(de);
    hipFree(df);
    for (int i = 0; i < n; i++) {
        Es[i] = e[i];
        Fs[i] = f[i];
    }
}
void sample_DistConstr(double lmin, double lmax, double kmin, double kmax, double flim, int n, double* xs, double* Es,
                       double* Fs) {
    (void)lmin; (void)lmax; (void)kmin; (void)kmax; (void)flim; (void)xs;
    unsupported("sample_DistConstr");
    for (int i = 0; i < n; i++) Es[i] = Fs[i] = 0;
}
void sample_evalPiAling(double k, double ang0, double r1, double r2, int n, double* angles, double* Es, double* Fs) {
    sample_angles_dev(0, k, ang0, r1, r2, n, angles, Es, Fs);
}
void sample_evalAngleCos(double k, double ang0, double r1, double r2, int n, double* angles, double* Es, double* Fs) {
    sample_angles_dev(1, k, ang0, r1, r2, n, angles, Es, Fs);
}
void sample_evalAngleCosHalf(double k, double ang0, double r1, double r2, int n, double* angles, double* Es,
                             double* Fs) {
    sample_angles_dev(2, k, ang0, r1, r2, n, angles, Es, Fs);
}
// kind 2 samples this library's pair term getLJQH (not the reference's
// addAtomicForceLJQ); the other kinds are not part of the path.
void sampleNonBond(int n, double* rs, double* Es, double* fs, int kind, double* REQi_, double* REQj_, double K,
                   double Rdamp) {
    (void)K;
    (void)Rdamp;
    World& W = world();
    if (n <= 0) return;
    if (kind != 2 || !W.ensure_device()) {
        unsupported("sampleNonBond(kind!=2)");
        for (int i = 0; i < n; i++) Es[i] = fs[i] = 0;
        return;
    }
    float4 req = make_float4((float)(REQi_[0] + REQj_[0]), (float)(REQi_[1] * REQj_[1]), (float)(REQi_[2] * REQj_[2]),
                             (float)(REQi_[3] * REQj_[3]));
    std::vector<float> r(n);
    std::vector<float2> o(n);
    for (int i = 0; i < n; i++) r[i] = (float)rs[i];
    float* dr;
    float2* dout;
    hipMalloc(&dr, n * 4);
    hipMalloc(&dout, n * 8);
    hipMemcpy(dr, r.data(), n * 4, hipMemcpyHostToDevice);
    fcw::hip_ok(fc_sample_pair(n, dr, req, dout), "sampleNonBond");
    hipMemcpy(o.data(), dout, n * 8, hipMemcpyDeviceToHost);
    hipFree(dr);
    hipFree(dout);
    for (int i = 0; i < n; i++) {
        fs[i] = o[i].x;
        Es[i] = o[i].y;
    }
}

// GridFF probe sampling (MMFFmulti_lib.cpp:168-250): kinds >= 10 sample this
// library's B-spline grid with a probe of element `atyp` (0=C,1=N,2=O,3=H) and
// charge Q; other kinds are not part of the path.
static void sample_grid_dev(int n, const double* p3, double* Es, double* f3, int ncomp, int atyp, double Q,
                            double K) {
    World& W = world();
    if (n <= 0 || !W.ensure_device() || W.grid_f4.empty()) {
        for (int i = 0; i < n; i++) Es[i] = 0;
        return;
    }
    if (!W.d_grid.p || W.d_grid.n != W.grid_f4.size()) {
        W.d_grid.upload(W.grid_f4, W.stream);
        W.sync();
    }
    const double R0[4] = {1.9255, 1.8300, 1.7500, 1.4430};
    const double E0[4] = {0.00455323095, 0.00299212319, 0.00260184625, 0.0020};
    int t = (atyp >= 0 && atyp < 4) ? atyp : 0;
    double alpha = K > 0 ? K : W.grid.alphaMorse;
    double e = std::exp(alpha * R0[t]);
    double cL = e * std::sqrt(E0[t]);
    float4 plq = make_float4((float)(e * cL), (float)cL, (float)Q, 0.f);
    std::vector<float4> pos(n), out(n);
    for (int i = 0; i < n; i++) pos[i] = make_float4((float)p3[i * 3], (float)p3[i * 3 + 1], (float)p3[i * 3 + 2], 0);
    float4 *dp, *dout;
    hipMalloc(&dp, n * 16);
    hipMalloc(&dout, n * 16);
    hipMemcpy(dp, pos.data(), n * 16, hipMemcpyHostToDevice);
    fcd::RunParams P = W.run_params(0, 0);
    P.grid.coefs = W.d_grid.p;
    fcw::hip_ok(fc_sample_grid(&P.grid, n, dp, plq, dout), "sampleSurf");
    hipMemcpy(out.data(), dout, n * 16, hipMemcpyDeviceToHost);
    hipFree(dp);
    hipFree(dout);
    for (int i = 0; i < n; i++) {
        Es[i] = out[i].w;
        if (ncomp == 1) f3[i] = out[i].z;
        else {
            f3[i * 3] = out[i].x;
            f3[i * 3 + 1] = out[i].y;
            f3[i * 3 + 2] = out[i].z;
        }
    }
}
void sampleSurf(char* name, int n, double* rs, double* Es, double* fs, int kind, int atyp, double Q, double K,
                double RQ, double* pos0_, bool bSave) {
    (void)RQ;
    (void)bSave;
    World& W = world();
    if (name) {
        fc::GridFFHost g;
        if (setup_surface(name, 0.1, g)) W.set_grid(g);
    }
    if (kind < 10) {
        unsupported("sampleSurf(kind<10)");
        for (int i = 0; i < n; i++) Es[i] = fs[i] = 0;
        return;
    }
    std::vector<double> p((size_t)n * 3);
    for (int i = 0; i < n; i++) {
        p[i * 3] = pos0_ ? pos0_[0] : 0;
        p[i * 3 + 1] = pos0_ ? pos0_[1] : 0;
        p[i * 3 + 2] = rs[i];
    }
    sample_grid_dev(n, p.data(), Es, fs, 1, atyp, Q, K);
}
void sampleSurf_vecs(char* name, int n, double* poss_, double* Es, double* fs_, int kind, int atyp, double Q,
                     double K, double RQ, double* pos0_, bool bSave) {
    (void)RQ;
    (void)bSave;
    (void)pos0_;
    World& W = world();
    if (name) {
        fc::GridFFHost g;
        if (setup_surface(name, 0.1, g)) W.set_grid(g);
    }
    if (kind < 10) {
        unsupported("sampleSurf_vecs(kind<10)");
        for (int i = 0; i < n; i++) Es[i] = 0;
        return;
    }
    sample_grid_dev(n, poss_, Es, fs_, 3, atyp, Q, K);
}

// ================================================================ additive
// Build a batch from CSR-packed host arrays (one entry per system in natoms /
// nnode; per-atom/per-node rows concatenated in system order).
int init_systems(int nSys, const int* natoms, const int* nnode, const double* apos, const double* REQs,
                 const int* neighs, const double* apars, const double* bLs, const double* bKs, const double* Ksp,
                 const double* Kpp, const int* elem) {
    World& W = world();
    std::vector<fc::MolSys> ms(nSys);
    size_t ia = 0, in = 0, iv = 0;
    for (int s = 0; s < nSys; s++) {
        fc::MolSys& m = ms[s];
        m.natoms = natoms[s];
        m.nnode = nnode[s];
        const int N = m.natoms, M = m.nnode, V = N + M;
        if (N <= 0 || M < 0 || M > N) return -1;
        m.apos.assign(apos + iv * 3, apos + (iv + V) * 3);
        m.REQs.assign(REQs + ia * 4, REQs + (ia + N) * 4);
        m.neighs.assign(neighs + ia * 4, neighs + (ia + N) * 4);
        for (int x : m.neighs)
            if (x >= N) return -1;
        m.apars.assign(apars + in * 4, apars + (in + M) * 4);
        m.bLs.assign(bLs + in * 4, bLs + (in + M) * 4);
        m.bKs.assign(bKs + in * 4, bKs + (in + M) * 4);
        m.Ksp.assign(Ksp + in * 4, Ksp + (in + M) * 4);
        m.Kpp.assign(Kpp + in * 4, Kpp + (in + M) * 4);
        if (elem) m.elem.assign(elem + ia, elem + ia + N);
        else m.elem.assign(N, 0);
        // pi-vectors not given (zero rows): initial ones from the geometry, as
        // MMFFsp3_loc::initPi + flipPis(1,1,1) do after building a molecule
        bool given = true;
        for (int k = 0; k < M && given; k++) {
            const double* p = &m.apos[(N + k) * 3];
            if (p[0] * p[0] + p[1] * p[1] + p[2] * p[2] < 1e-12) given = false;
        }
        if (!given) fc::init_pi(N, M, m.apos.data(), m.neighs.data(), m.Ksp.data(), m.Kpp.data(), true,
                                m.apos.data() + (size_t)N * 3);
        ia += N;
        in += M;
        iv += V;
    }
    W.bGridFF = false;  // a new batch has no substrate until one is loaded
    W.has_templ = false;
    if (!W.set_systems(std::move(ms))) return -1;
    if (!finish_init(W)) return -1;
    return W.ntotA;
}

// Synthetic workloads.
//  kind 0: one molecule type of nmin atoms (seed seed0), nSys replicas each
//          perturbed with seed seed0+1+i (uniform batch)
//  kind 1: nSys independent molecules, natoms ~ U[nmin,nmax], seed seed0+i
//  kind 2: nSys pairs of molecules, each ~ U[nmin,nmax] atoms, random relative
//          orientation, 4-8 A apart
// surf_name (may be NULL) places every system 3.0 +- 0.3 A above the substrate
// and loads/creates its GridFF.
int init_workload(int kind, int nSys, long long seed0, int nmin, int nmax, double sigma, const char* surf_name,
                  double gridStep) {
    World& W = world();
    std::vector<fc::MolSys> ms(nSys);
    fc::GridFFHost g;
    bool surf = surf_name && surf_name[0];
    if (surf && !setup_surface(surf_name, gridStep > 0 ? gridStep : 0.1, g)) return -1;
    const double ax = g.nx * g.step[0], by = g.ny * g.step[1];
    std::atomic<bool> fail{false};
    if (kind == 0) {
        fc::MolSys tmpl;
        if (!fc::gen_molecule(nmin, (uint64_t)seed0, tmpl)) return -1;
        for (int i = 0; i < nSys; i++) {
            ms[i] = tmpl;
            fc::perturb_positions(ms[i], (uint64_t)(seed0 + 1 + i), sigma);
        }
    } else {
        const int nth = 16;
        std::vector<std::thread> th;
        for (int t = 0; t < nth; t++)
            th.emplace_back([&, t]() {
                for (int i = t; i < nSys; i += nth)
                    if (!fc::make_workload_system(kind, i, seed0, nmin, nmax, sigma, surf, g.pos0, ax, by, ms[i]))
                        fail = true;
            });
        for (auto& x : th) x.join();
    }
    if (fail) return -1;
    W.bGridFF = false;
    W.has_templ = false;
    if (!W.set_systems(std::move(ms))) return -1;
    if (surf) W.set_grid(g);
    if (!finish_init(W)) return -1;
    return W.ntotA;
}

int init_pi_vectors(int natoms, int nnode, const double* apos, const int* neighs, const double* Ksp,
                    const double* Kpp, int flip, double* pipos) {
    if (natoms <= 0 || nnode < 0 || nnode > natoms || !apos || !neighs || !Ksp || !Kpp || !pipos) return -1;
    for (int i = 0; i < natoms * 4; i++)
        if (neighs[i] >= natoms) return -1;
    fc::init_pi(natoms, nnode, apos, neighs, Ksp, Kpp, flip != 0, pipos);
    return 0;
}

int attach_surface(const char* surf_name, double gridStep) {
    World& W = world();
    if (!surf_name || !surf_name[0] || W.nSys == 0) return -1;
    fc::GridFFHost g;
    if (!setup_surface(surf_name, gridStep > 0 ? gridStep : 0.1, g)) return -1;
    return W.set_grid(g) ? 0 : -1;
}

// The substrate of surf_name as explicit atoms with images -nPBC..nPBC of its
// lattice (GridFF::evalMorsePBC, the reference's CPU path without a grid);
// replaces the GridFF.  surf_name = null removes the substrate atoms.
int attach_surface_atoms(const char* surf_name, const int* nPBC3) {
    World& W = world();
    if (W.nSys == 0) return -1;
    fc::Surface S;
    if (surf_name && !load_surface_atoms(surf_name, S)) return -1;
    const int np0[3] = {1, 1, 0};
    return W.set_surface_atoms(S, nPBC3 ? nPBC3 : np0) ? 0 : -1;
}

// Substrate atoms (double): those of surf_name read on the host (no device
// work; nPBC3 untouched), or with surf_name = null the ones in use.  Positions
// (n x 3), REQs (n x 4), lattice rows (9) and image range (3), at most maxn
// atoms.  Returns their number, -1 if surf_name cannot be read.
int get_surface_atoms(const char* surf_name, double* apos, double* REQs, double* lvec9, int* nPBC3, int maxn) {
    const World& W = world();
    fc::Surface loaded;
    if (surf_name && !load_surface_atoms(surf_name, loaded)) return -1;
    const fc::Surface& S = surf_name ? loaded : W.surf_host;
    const int n = (surf_name || !W.surf_pos.empty()) ? S.natoms() : 0;
    const int m = std::min(n, std::max(maxn, 0));
    if (apos) std::memcpy(apos, S.apos.data(), sizeof(double) * 3 * m);
    if (REQs) std::memcpy(REQs, S.REQs.data(), sizeof(double) * 4 * m);
    if (lvec9) std::memcpy(lvec9, S.lvec, sizeof(double) * 9);
    if (nPBC3 && !surf_name)
        for (int k = 0; k < 3; k++) nPBC3[k] = W.surf_nPBC[k];
    return n;
}

// Relax the whole batch on the device.
//   path 0: persistent kernel; 1: split-phase kernels
//   restore_start: restart from the positions of the last init/upload
//   bDownload: copy atoms/forces/velocities back into the host mirrors
// Returns the largest number of FIRE steps taken by any system, -1 on error.
int relax(int nstepMax, double Fconv, int path, bool restore_start, bool bDownload) {
    World& W = world();
    if (!W.dev_ready || W.nSys == 0) return -1;
    int n = (path == 1) ? W.relax_split(nstepMax, Fconv, restore_start) : W.relax_persistent(nstepMax, Fconv, restore_start);
    if (n >= 0 && bDownload && !W.download_all(true, true)) return -1;
    return n;
}

// relax() path 0 with no device allocation: the persistent kernel's scratch
// (queue counters and order, network activation slots, GridFF window cache) in
// the library's static device array of FC_NOALLOC_SCRATCH_BYTES.  Rejects
// (returns -1 with a message) batches that need more scratch or a path that
// allocates.  Calls that overlap on one device share its scratch and race.
int relax_noalloc(int nstepMax, double Fconv, bool restore_start, bool bDownload) {
    World& W = world();
    if (!W.dev_ready || W.nSys == 0) return -1;
    int n = W.relax_noalloc(nstepMax, Fconv, restore_start);
    if (n >= 0 && bDownload && !W.download_all(true, true)) return -1;
    return n;
}

// Global-optimisation exploring (MolWorld_sp3_multi::updateMultiExploring,
// MolWorld_sp3_multi.h:770-833, with the Langevin drive of updateAtomsMMFFf4,
// relax_multi.cl:1160-1178): every system alternates FIRE relaxation and
// Langevin exploring (nExplore steps at temperature T, damping gamma) for
// nstepMax steps from its current state, recording up to kmax minima (energy,
// and positions when bPositions).  Returns the number of minima found over all
// systems, -1 on error.
int explore(int nstepMax, double Fconv, double T, double gamma, int nExplore, int nRelax, long long seed, int kmax,
            bool bPositions) {
    World& W = world();
    if (!W.dev_ready || W.nSys == 0) return -1;
    int n = W.explore(nstepMax, Fconv, T, gamma, nExplore, nRelax, (unsigned)seed, kmax, bPositions);
    if (n >= 0) W.download_all(true, true);
    return n;
}

// Minima recorded for system isys by the last explore(): energies into Es and,
// if recorded and poss is given, positions (natoms+nnode rows of x,y,z per
// minimum) into poss; at most maxn.  Returns the number of minima found (may
// exceed the number recorded), -1 on error.
int get_minima(int isys, double* Es, double* poss, int maxn) {
    World& W = world();
    if (isys < 0 || isys >= W.nSys || W.nmin.size() != (size_t)W.nSys) return -1;
    const int found = W.nmin[isys];
    const int nrec = std::min(std::min(found, W.xkmax), std::max(maxn, 0));
    const fcd::SysInfo& I = W.info[isys];
    const int V = I.natoms + I.nnode;
    for (int m = 0; m < nrec; m++) {
        if (Es) Es[m] = W.minE[(size_t)isys * W.xkmax + m];
        if (poss && W.minpos_kept)
            for (int v = 0; v < V; v++) {
                const float4 p = W.minpos[(size_t)m * W.ntotV + I.v0 + v];
                poss[((size_t)m * V + v) * 3] = p.x;
                poss[((size_t)m * V + v) * 3 + 1] = p.y;
                poss[((size_t)m * V + v) * 3 + 2] = p.z;
            }
    }
    return found;
}

// Minima database of the exploring mode (MolecularDatabase, filled by explore()
// and by run() after init(nExplore > 0) while setSwitches_multi(...,
// bSaveToDatabase = 1) is on; the reference appends new members to gopt.xyz).
// i < 0: the number of members.  Else member i's atom count, with its energy,
// system, occurrence count and positions (natoms x 3, at most maxatoms rows)
// written where the pointers are not null; -1 when i is out of range.
int get_database_member(int i, double* E, int* isys, int* occurrence, double* pos, int maxatoms) {
    const fc::MinimaDB& D = world().db;
    if (i < 0) return (int)D.members.size();
    if (i >= (int)D.members.size()) return -1;
    const fc::MinimaDB::Member& m = D.members[i];
    if (E) *E = m.E;
    if (isys) *isys = m.isys;
    if (occurrence) *occurrence = m.occurrence;
    if (pos) std::memcpy(pos, m.pos.data(), sizeof(double) * 3 * std::min(m.natoms, std::max(maxatoms, 0)));
    return m.natoms;
}
// Offer a structure of n atoms (positions n x 3, a type label per atom) to
// the database: the index of the member it duplicates, or -1 when it became a
// new member.  Host only.
int database_add(int n, const double* pos, const int* types, double E) {
    if (n < 1 || !pos || !types) return -2;
    return world().db.add_if_new(n, pos, types, E, -1);
}
void clear_database() { world().db.clear(); }
// Descriptor of a structure (sqrt of the inertia tensor's eigenvalues about the
// centroid, ascending, then n; the eigenvalues only for n >= 3): its length.
int database_descriptor(int n, const double* pos, double* d4) { return fc::db_descriptor(n, pos, d4); }
// Distance the database compares with 0.1 n: sqrt of the sum over y's atoms of
// the squared distance to the nearest same-type atom of x, after the optimal
// superposition of the centred x on the centred y.
double database_distance(int n, const double* x, const int* tx, const double* y, const int* ty) {
    return fc::db_distance(n, x, tx, y, ty);
}
// The members as xyz frames (element symbols of the type labels C, N, O, H
// codes 0-3, else X), comment "# <index> E <energy> occurrence <count> isys <s>".
int save_database_xyz(const char* fname) {
    const fc::MinimaDB& D = world().db;
    FILE* f = std::fopen(fname, "w");
    if (!f) return -1;
    static const char* sym[4] = {"C", "N", "O", "H"};
    for (size_t i = 0; i < D.members.size(); i++) {
        const fc::MinimaDB::Member& m = D.members[i];
        std::fprintf(f, "%d\n# %zu E %.10g occurrence %d isys %d\n", m.natoms, i, m.E, m.occurrence, m.isys);
        for (int a = 0; a < m.natoms; a++) {
            const int t = m.types[a];
            std::fprintf(f, "%s %.8f %.8f %.8f\n", (t >= 0 && t < 4) ? sym[t] : "X", m.pos[a * 3], m.pos[a * 3 + 1],
                         m.pos[a * 3 + 2]);
        }
    }
    std::fclose(f);
    return (int)D.members.size();
}

// Periodic molecules (MMFFsp3_loc with bPBC: image sums of the non-bonded
// terms, bonded neighbours at their nearest image, MMFFsp3_loc.h:325-329,
// NBFF.h:575-620): lattice rows lvec9 (a, b, c) and image range nPBC3 (null =
// keep the system's, or init's) of system isys, or of all systems if isys < 0;
// lvec9 = null makes them vacuum molecules again.  Periodic systems are relaxed
// by the split kernels.  0 on success, -1 on error.
int set_lattice(int isys, const double* lvec9, const int* nPBC3) {
    World& W = world();
    if (W.nSys == 0 || isys >= W.nSys) return -1;
    const int* np = nPBC3 ? nPBC3 : W.nPBC_init;
    return W.set_lattice(isys, lvec9, np) ? 0 : -1;
}

double get_last_ms() { return world().last_ms; }

int set_cost_model(int feedback, double c0, double c1) {
    World& W = world();
    if (!std::isfinite(c0) || !std::isfinite(c1)) return -1;
    W.cost_feedback = feedback != 0;
    W.cost_c0 = c0;
    W.cost_c1 = c1;
    return 0;
}

int set_class_mask(int mask) {
    World& W = world();
    if (mask < 0 || mask > 0xF) return -1;
    const int old = W.class_mask;
    W.class_mask = mask | 0x8;
    if (W.dev_ready && W.nSys > 0) W.plan_classes();
    return old;
}

// Waits for all work on the device and returns 0, or the HIP error code of a
// failure that earlier calls have not reported (message on stderr).
int device_status() {
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) std::fprintf(stderr, "[libMMFFmulti_lib] device status: %s\n", hipGetErrorString(e));
    return (int)e;
}

int set_record_times(int on) {
    World& W = world();
    const int old = W.record_times ? 1 : 0;
    W.record_times = on != 0;
    return old;
}

// Per-system start and end of the last persistent relaxation, in microseconds
// after the earliest start (device wall clock), into out[2*s], out[2*s+1].
int get_system_times(double* out) {
    World& W = world();
    if (!W.record_times || !W.d_systime.p || W.nSys == 0) return -1;
    std::vector<long long> t((size_t)2 * W.nSys);
    W.d_systime.download(t, W.stream);
    if (!fcw::hip_ok(hipStreamSynchronize(W.stream), "system times")) return -1;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0) != hipSuccess || khz <= 0) return -1;
    long long t0 = t[0];
    for (int s = 0; s < W.nSys; s++) t0 = std::min(t0, t[2 * s]);
    for (size_t i = 0; i < t.size(); i++) out[i] = (double)(t[i] - t0) * 1000.0 / khz;
    return 0;
}

// Size classes of the persistent kernel (one launch each): per class
// {block, nsys, nmax, mmax, ngroups, lds_bytes} into info6[6*k] and the device
// time of its launch in the last relax into ms[k].  Returns the class count.
int get_class_info(int* info6, double* ms, int maxc) {
    World& W = world();
    int n = std::min((int)W.classes.size(), maxc);
    for (int c = 0; c < n; c++) {
        const fcw::SizeClass& C = W.classes[c];
        info6[6 * c + 0] = C.block;
        info6[6 * c + 1] = (int)C.order.size();
        info6[6 * c + 2] = C.nmax;
        info6[6 * c + 3] = C.mmax;
        info6[6 * c + 4] = C.ngroups;
        info6[6 * c + 5] = (int)C.lds;
        ms[c] = C.last_ms;
    }
    return n;
}
// System indices of class c in queue order (largest first); returns count.
int get_class_order(int c, int* out, int maxn) {
    World& W = world();
    if (c < 0 || c >= (int)W.classes.size()) return 0;
    int n = std::min((int)W.classes[c].order.size(), maxn);
    for (int i = 0; i < n; i++) out[i] = W.classes[c].order[i];
    return n;
}

// Evaluate forces and energies of every system at the current positions
// (split kernels, no move); results in gpu_aforces / sys_E after download.
int eval_forces() {
    World& W = world();
    if (!W.eval_forces()) return -1;
    return W.download_all(true, false) ? 0 : -1;
}

void set_conv_mode(int mode) { world().conv_mode = mode; }

// Per-vec force clamp for the FIRE dynamics (the reference GPU update kernel
// clamps |f| to 10 eV/A, relax_multi.cl:980-983); 0 disables it (the CPU
// path's behaviour).  Convergence is always tested on the unclamped force.
void set_force_limit(double Flimit) { world().Flimit = (float)Flimit; }

// FIRE force preconditioner of the persistent path: mode 0 off (the
// reference dynamics), 1 rigid-group projections with constant gains
// (gstar per bonded star, grigid per molecule), 2 gains predicted by the
// proposal network every `every` steps with gstar/grigid as ceilings.
// Returns 0, or -1 for an invalid mode or gain.
int set_precond(int mode, double gstar, double grigid) {
    if (mode < 0 || mode > 2 || !(gstar >= 0.0) || !(grigid >= 0.0)) {
        std::fprintf(stderr, "[libMMFFmulti_lib] set_precond: invalid mode %d or gains %g %g\n", mode, gstar, grigid);
        return -1;
    }
    World& W = world();
    if (mode == 2 && !W.d_mlpW.p) {
        std::fprintf(stderr, "[libMMFFmulti_lib] set_precond: mode 2 needs proposal weights (set_proposal_weights)\n");
        return -1;
    }
    W.pc_mode = mode;
    W.pc_gstar = (float)gstar;
    W.pc_grigid = (float)grigid;
    if (W.dev_ready && W.nSys > 0) W.plan_classes();
    return 0;
}

// Exact forces of every system (download into gpu_aforces) and the
// preconditioned force of the current set_precond gains (download into the
// additive view gpu_pcforces).  0 on success.
int eval_precond() {
    World& W = world();
    return W.eval_precond() ? 0 : -1;
}

// ---- proposal network (no counterpart in the reference) ----------------
// Parameters in the packing order W1[64][16] b1[64] W2[64][128] b2[64]
// W3[64][128] b3[64] W4[2][64] b4[2] (FP32); dense weights are used in bf16.
int get_proposal_nparam() { return 64 * 16 + 64 + 2 * (64 * 128 + 64) + 128 + 2; }
int set_proposal_weights(const float* theta, int n) {
    World& W = world();
    if (!theta || !W.set_mlp(theta, n)) return -1;
    W.theta_dev = false;  // training restarts from these parameters
    return 0;
}
int get_proposal_weights(float* theta, int n) {
    World& W = world();
    if (!W.download_theta()) return -1;
    if ((int)W.mlp_theta.size() != n || !theta) return -1;
    std::memcpy(theta, W.mlp_theta.data(), sizeof(float) * n);
    return 0;
}
// steps between network evaluations inside the relaxation (>= 1)
int set_proposal_every(int every) {
    if (every < 1) return -1;
    world().pc_every = every;
    return 0;
}
// FIRE fallback of the learned form: a system whose max|F| at `windows`
// proposal points in a row is no lower than its best one so far continues
// with plain FIRE (zero gains, no further network evaluations).  0 = never.
int set_proposal_stall(int windows) {
    if (windows < 0) return -1;
    world().pc_stall = windows;
    return 0;
}
// Per system: 1 when the learned form fell back to plain FIRE in the last
// relaxation, else 0 (learned form only).  0 on success.
int get_proposal_fallback(int* out) {
    World& W = world();
    if (!W.dev_ready || !out) return -1;
    std::vector<float4> st(W.nSys);
    W.d_stall.download(st, W.stream);
    if (!fcw::hip_ok(hipStreamSynchronize(W.stream), "fallback flags")) return -1;
    for (int s = 0; s < W.nSys; s++) out[s] = st[s].z != 0.f ? 1 : 0;
    return 0;
}
// Record trajectory snapshots during the next relax() calls: positions of
// every vec at steps 0, every, 2*every, ... (n snapshots at most).  n = 0
// stops recording.  0 on success.
int record_snapshots(int every, int n) { return world().record_snapshots(every, n) ? 0 : -1; }
// Load snapshot idx[s] of every system s into its positions.
int load_snapshots(const int* idx) {
    World& W = world();
    std::vector<int> v(idx, idx + W.nSys);
    return W.load_snapshots(v) ? 0 : -1;
}
// Forward pass of the proposal network at the current positions (exact
// forces are evaluated first).  out: [ntotA][2] outputs; feats (optional):
// [ntotA][16] the inputs as fed to the matrix cores (bf16 values as float).
int proposal_forward(float* out, float* feats) {
    World& W = world();
    if (!W.prop_prepare() || !W.prop_forward()) return -1;
    std::vector<float> o((size_t)W.ntotA * 2);
    W.d_out.download(o, W.stream);
    std::vector<unsigned short> x;
    if (feats) {
        x.resize((size_t)W.ntotA * 16);
        W.d_X.download(x, W.stream);
    }
    if (!fcw::hip_ok(hipStreamSynchronize(W.stream), "proposal_forward")) return -1;
    if (out) std::memcpy(out, o.data(), sizeof(float) * o.size());
    if (feats)
        for (size_t i = 0; i < x.size(); i++) {
            const uint32_t u = ((uint32_t)x[i]) << 16;
            std::memcpy(feats + i, &u, 4);
        }
    return 0;
}
// Training loss of one preconditioned step from the current positions after
// proposal_forward: per-system loss (E(x0 + eta*F~) - E(x0)) / N into loss
// [nSys], dloss/doutputs into dout [ntotA][2] (either may be NULL).
int proposal_loss_grad(double eta, double dmax, float* dout, float* loss) {
    World& W = world();
    if (!W.prop_loss_grad((float)eta, (float)dmax)) return -1;
    std::vector<float> d((size_t)W.ntotA * 2), l(W.nSys);
    W.d_dout.download(d, W.stream);
    W.d_loss.download(l, W.stream);
    if (!fcw::hip_ok(hipStreamSynchronize(W.stream), "proposal_loss_grad")) return -1;
    if (dout) std::memcpy(dout, d.data(), sizeof(float) * d.size());
    if (loss) std::memcpy(loss, l.data(), sizeof(float) * l.size());
    return 0;
}
// Backward pass after proposal_forward for the given dloss/doutputs
// ([ntotA][2]); grad [nparam] receives the mean over systems of the
// per-system parameter gradients.
int proposal_backward(const float* dout, float* grad) {
    World& W = world();
    if (!W.d_dout.p || !dout) return -1;
    std::vector<float> d(dout, dout + (size_t)W.ntotA * 2);
    W.d_dout.upload(d, W.stream);
    if (!W.prop_backward()) return -1;
    std::vector<float> g(get_proposal_nparam());
    W.d_grad.download(g, W.stream);
    if (!fcw::hip_ok(hipStreamSynchronize(W.stream), "proposal_backward")) return -1;
    if (grad) std::memcpy(grad, g.data(), sizeof(float) * g.size());
    return 0;
}
// Train the proposal network on the recorded snapshots: iters steps of Adam
// (learning rate lr), each on one snapshot per system drawn uniformly from
// the system's valid snapshots (counts from the last relax).  loss_hist
// (optional, [iters]) receives the mean loss of every step.
int train_proposal(int iters, double lr, double eta, double dmax, long long seed, float* loss_hist) {
    World& W = world();
    if (!W.snap_on || W.nsnap < 1 || iters < 0) return -1;
    std::vector<int> cnt(W.nSys);
    for (int s = 0; s < W.nSys; s++)
        cnt[s] = std::max(1, std::min(W.nsnap, W.nsteps[s] / W.snap_every + 1));
    fc::Rng rng((uint64_t)seed);
    std::vector<int> idx(W.nSys);
    std::vector<float> lm(1);
    for (int it = 0; it < iters; it++) {
        for (int s = 0; s < W.nSys; s++) idx[s] = (int)(rng.next() % (uint64_t)cnt[s]);
        if (!W.load_snapshots(idx) || !W.prop_prepare() || !W.prop_forward() ||
            !W.prop_loss_grad((float)eta, (float)dmax) || !W.prop_backward() || !W.prop_adam((float)lr))
            return -1;
        if (loss_hist) {
            W.d_lossmean.download(lm, W.stream);
            if (!fcw::hip_ok(hipStreamSynchronize(W.stream), "train")) return -1;
            loss_hist[it] = lm[0];
        }
    }
    W.restore_positions();
    return W.download_theta() ? 0 : -1;
}

// 2 = evaluate the proposal network inside the persistent relaxation kernel
// with its second activation buffer in global memory (default; leaves the
// kernel's LDS footprint, and so its occupancy, unchanged), 1 = the same with
// that buffer in LDS where the class's LDS budget allows, 0 = chunked
// relaunches with a separate network kernel between chunks (graph replay).
// Returns the previous setting.
int set_learned_inkernel(int on) {
    World& W = world();
    const int old = W.learned_inkernel ? (W.h2lds_allowed ? 1 : 2) : 0;
    W.learned_inkernel = on != 0;
    W.h2lds_allowed = on != 2;
    return old;
}

// Chunk iterations of the last learned relaxation; 0 after other paths.
int get_last_chunks() { return world().last_chunks; }
// 1 = replay the learned-relaxation chunk loop from a captured graph
// (default), 0 = launch every chunk directly.  Returns the previous setting.
int set_graphs(int on) {
    World& W = world();
    const int old = W.use_graphs ? 1 : 0;
    W.use_graphs = on != 0;
    return old;
}

// Workgroup scheduling of the persistent path: 0 = device slots split between
// the size classes by estimated time with work stealing, 1 = every class
// launched at full device width (the dispatcher interleaves them), no
// stealing.  Returns the previous mode, or -1 for an invalid mode.
int set_sched(int mode) {
    World& W = world();
    if (mode < 0 || mode > 2) return -1;
    const int old = W.sched_mode;
    W.sched_mode = mode;
    if (W.dev_ready && W.nSys > 0) W.plan_classes();
    return old;
}

// nSys, ntotA, ntotM, ntotV, maxN, maxM, grid nx, ny, nz
void get_batch_sizes(int* out) {
    World& W = world();
    out[0] = W.nSys;
    out[1] = W.ntotA;
    out[2] = W.ntotM;
    out[3] = W.ntotV;
    out[4] = W.maxN;
    out[5] = W.maxM;
    out[6] = W.grid.nx;
    out[7] = W.grid.ny;
    out[8] = W.grid.nz;
}

// Grid geometry: pos0[3], step[3], alphaMorse; returns the coefficient count
int get_grid_info(double* out7) {
    World& W = world();
    for (int k = 0; k < 3; k++) {
        out7[k] = W.grid.pos0[k];
        out7[3 + k] = W.grid.step[k];
    }
    out7[6] = W.grid.alphaMorse;
    return (int)W.grid.coefs.size();
}
double* get_grid_coefs() { return world().grid.coefs.data(); }

// Host-only helpers (no device): generator and topology, for inspection.
int gen_molecule_arrays(int natoms, long long seed, double sigma, long long pseed, int* nnode_out, double* apos,
                        double* REQs, int* neighs, double* apars, double* bLs, double* bKs, double* Ksp, double* Kpp,
                        int* elem) {
    fc::MolSys m;
    if (!fc::gen_molecule(natoms, (uint64_t)seed, m)) return -1;
    if (sigma > 0) fc::perturb_positions(m, (uint64_t)pseed, sigma);
    *nnode_out = m.nnode;
    if (apos) std::memcpy(apos, m.apos.data(), m.apos.size() * 8);
    if (REQs) std::memcpy(REQs, m.REQs.data(), m.REQs.size() * 8);
    if (neighs) std::memcpy(neighs, m.neighs.data(), m.neighs.size() * 4);
    if (apars) std::memcpy(apars, m.apars.data(), m.apars.size() * 8);
    if (bLs) std::memcpy(bLs, m.bLs.data(), m.bLs.size() * 8);
    if (bKs) std::memcpy(bKs, m.bKs.data(), m.bKs.size() * 8);
    if (Ksp) std::memcpy(Ksp, m.Ksp.data(), m.Ksp.size() * 8);
    if (Kpp) std::memcpy(Kpp, m.Kpp.data(), m.Kpp.size() * 8);
    if (elem) std::memcpy(elem, m.elem.data(), m.elem.size() * 4);
    return m.nnode;
}
// One system of init_workload (kind 1 or 2), generated on the host only.
// pos0/cell: grid origin and lateral cell of the substrate (NULL: vacuum).
int gen_workload_arrays(int kind, int i, long long seed0, int nmin, int nmax, double sigma, const double* pos0,
                        const double* cell, int* natoms_out, int* nnode_out, double* apos, double* REQs, int* neighs,
                        double* apars, double* bLs, double* bKs, double* Ksp, double* Kpp, int* elem) {
    fc::MolSys m;
    const double z[3] = {0, 0, 0};
    if (!fc::make_workload_system(kind, i, seed0, nmin, nmax, sigma, pos0 != nullptr, pos0 ? pos0 : z,
                                  cell ? cell[0] : 0, cell ? cell[1] : 0, m))
        return -1;
    *natoms_out = m.natoms;
    *nnode_out = m.nnode;
    if (apos) std::memcpy(apos, m.apos.data(), m.apos.size() * 8);
    if (REQs) std::memcpy(REQs, m.REQs.data(), m.REQs.size() * 8);
    if (neighs) std::memcpy(neighs, m.neighs.data(), m.neighs.size() * 4);
    if (apars) std::memcpy(apars, m.apars.data(), m.apars.size() * 8);
    if (bLs) std::memcpy(bLs, m.bLs.data(), m.bLs.size() * 8);
    if (bKs) std::memcpy(bKs, m.bKs.data(), m.bKs.size() * 8);
    if (Ksp) std::memcpy(Ksp, m.Ksp.data(), m.Ksp.size() * 8);
    if (Kpp) std::memcpy(Kpp, m.Kpp.data(), m.Kpp.size() * 8);
    if (elem) std::memcpy(elem, m.elem.data(), m.elem.size() * 4);
    return m.natoms;
}
// Typed molecule of an .xyz file on the host (the builder of init(xyz_name)):
// tables from the given ElementTypes/AtomTypes/BondTypes files.  Outputs in the
// MMFFsp3_loc layout (apos (natoms+nnode) x 3 with the pi-vectors, REQs x 4,
// neighs x 4, apars/bLs/bKs/Ksp/Kpp nnode x 4, elem), the lattice and the
// periodic flag.  Returns natoms, -1 on error (message on stderr), -2 if the
// molecule has more than maxn atoms.
int build_xyz_arrays(const char* xyz, const char* fElem, const char* fAtom, const char* fBond, int maxn,
                     int* nnode_out, double* apos, double* REQs, int* neighs, double* apars, double* bLs, double* bKs,
                     double* Ksp, double* Kpp, int* elem, double* lvec9, int* periodic) {
    fc::TypeTables T;
    std::string err;
    if (!T.load(fElem, fAtom, fBond, err)) {
        std::fprintf(stderr, "[libMMFFmulti_lib] build_xyz_arrays: %s\n", err.c_str());
        return -1;
    }
    fc::MolSys m;
    bool pbc = false;
    double lv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (!fc::build_from_xyz(xyz, T, pbc, lv, m, err)) {
        std::fprintf(stderr, "[libMMFFmulti_lib] build_xyz_arrays(%s): %s\n", xyz, err.c_str());
        return -1;
    }
    if (m.natoms > maxn) return -2;
    *nnode_out = m.nnode;
    std::memcpy(apos, m.apos.data(), m.apos.size() * 8);
    std::memcpy(REQs, m.REQs.data(), m.REQs.size() * 8);
    std::memcpy(neighs, m.neighs.data(), m.neighs.size() * 4);
    std::memcpy(apars, m.apars.data(), m.apars.size() * 8);
    std::memcpy(bLs, m.bLs.data(), m.bLs.size() * 8);
    std::memcpy(bKs, m.bKs.data(), m.bKs.size() * 8);
    std::memcpy(Ksp, m.Ksp.data(), m.Ksp.size() * 8);
    std::memcpy(Kpp, m.Kpp.data(), m.Kpp.size() * 8);
    std::memcpy(elem, m.elem.data(), m.elem.size() * 4);
    std::memcpy(lvec9, lv, sizeof(lv));
    *periodic = pbc ? 1 : 0;
    return m.natoms;
}

// Image shifts (evalPBCshifts order) and bonded-neighbour images
// (makeNeighCells) of a periodic molecule; returns npbc.
int make_neigh_cells(int natoms, const double* apos, const int* neighs, const double* lvec9, const int* nPBC3,
                     double* shifts_out, int* neighCell_out) {
    std::vector<double> sh;
    const int npbc = fc::make_pbc_shifts(lvec9, nPBC3, sh);
    if (shifts_out) std::memcpy(shifts_out, sh.data(), sh.size() * sizeof(double));
    if (neighCell_out) fc::make_neigh_cells(natoms, apos, neighs, npbc, sh.data(), neighCell_out);
    return npbc;
}
int make_back_neighs(int natoms, int nnode, const int* neighs, int* bkneighs) {
    return fc::make_back_neighs(natoms, nnode, neighs, bkneighs) ? 0 : -1;
}
int make_exclusions(int natoms, int nnode, const int* neighs, int subtract13, unsigned* mask_out, int* W_out,
                    int* ncorr_out) {
    std::vector<uint32_t> mask;
    std::vector<fc::PairCorr> corr;
    int W = 0;
    fc::make_exclusions(natoms, nnode, neighs, subtract13 != 0, mask, W, corr);
    if (mask_out) std::memcpy(mask_out, mask.data(), mask.size() * 4);
    *W_out = W;
    *ncorr_out = (int)corr.size();
    return 0;
}
// NaCl slab grid (direct sums + B-spline fit) without touching the device.
// GridFF node potentials (Pauli, London, Coulomb) of a substrate at npts points
// (host FP64 evaluation of the generator's definition).
int gridff_point_values(int natoms, const double* apos, const double* REQs, const double* lvec9, double alpha,
                        int npts, const double* pts, double* out) {
    if (natoms <= 0 || npts < 0) return -1;
    fc::Surface S;
    S.apos.assign(apos, apos + natoms * 3);
    S.REQs.assign(REQs, REQs + natoms * 4);
    std::memcpy(S.lvec, lvec9, sizeof(S.lvec));
    int nmx, nmy, ncx, ncy;
    fc::gridff_images(S, nmx, nmy, ncx, ncy);
    for (int i = 0; i < npts; i++) fc::gridff_point(S, pts + i * 3, alpha, nmx, nmy, ncx, ncy, out + i * 3);
    return 0;
}

// Node potentials of the synthetic NaCl(001) GridFF grid (origin at the top
// atomic plane, as make_nacl_gridff) computed by the device kernel (device=1)
// or the host (device=0), (nx,ny,nz,3) z fastest.
int gridff_nacl_nodes(int nx, int ny, int nz, double dz, double Rdamp, int device, double* out) {
    fc::Surface S;
    fc::make_nacl_slab(5.64, 2, 0.9, S);
    S.Rdamp = Rdamp;
    fc::GridFFHost g;
    fc::gridff_geometry(S, nx, ny, nz, dz, 0.0, 1.5, g);
    if (device) {
        if (!world().ensure_device()) return -1;
        return fcw::hip_ok(fc_gridff_nodes_device(&S, &g, out), "GridFF node potentials") ? 0 : -1;
    }
    std::vector<double> V;
    fc::gridff_nodes_host(S, g, 16, V);
    std::memcpy(out, V.data(), V.size() * sizeof(double));
    return 0;
}

// Exact cubic B-spline coefficients of node values V (nx,ny,nz,3), periodic in
// x,y and open in z, as used for every GridFF this library builds.
int gridff_fit_values(int nx, int ny, int nz, const double* V, double* coefs) {
    if (nx < 4 || ny < 4 || nz < 4) return -1;
    fc::GridFFHost g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    std::vector<double> v(V, V + (size_t)nx * ny * nz * 3);
    fc::gridff_fit(g, v, 8);
    std::memcpy(coefs, g.coefs.data(), g.coefs.size() * sizeof(double));
    return 0;
}

int make_nacl_gridff(int nx, int ny, int nz, double dz, double Rdamp, int nthreads, double* coefs_out,
                     double* geom_out7) {
    fc::Surface S;
    fc::make_nacl_slab(5.64, 2, 0.9, S);
    S.Rdamp = Rdamp;
    fc::GridFFHost g;
    fc::gridff_geometry(S, nx, ny, nz, dz, 0.0, 1.5, g);
    std::vector<double> V;
    fc::gridff_nodes_host(S, g, nthreads, V);
    fc::gridff_fit(g, V, nthreads);
    if (coefs_out) std::memcpy(coefs_out, g.coefs.data(), g.coefs.size() * 8);
    for (int k = 0; k < 3; k++) {
        geom_out7[k] = g.pos0[k];
        geom_out7[3 + k] = g.step[k];
    }
    geom_out7[6] = g.alphaMorse;
    return 0;
}

}  // extern "C"

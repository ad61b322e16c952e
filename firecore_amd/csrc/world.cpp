// Host world: CSR packing of uneven batches, device buffers and the device
// entry points (persistent relaxation, split-phase relaxation, force
// evaluation).  Packing conventions follow MolWorld_sp3_multi::pack_system
// (MolWorld_sp3_multi.h:514-592): neighs/neighCell are system-local,
// bkNeighs are absolute indices into fneigh (+n0*8 for atoms, +n0*8+4*nnode
// for pi rows, datatypes_utils.h:19 copy_add).
#include "world.h"

#include "precond.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

extern "C" hipError_t fc_launch_relax_persistent(int block, int hb, int ngroups, size_t lds_bytes, hipStream_t stream,
                                                 const fcd::BatchDev* B, const fcd::RunParams* P, const int* order,
                                                 const int* qinfo, int* counters, int cls, int nmax, int mmax,
                                                 int wmax, int steal, unsigned short* hslot, int h2lds);
extern "C" int fc_hslot_elems(int nmax);
extern "C" hipError_t fc_noalloc_scratch(void** ptr);
extern "C" hipError_t fc_launch_zero_int(hipStream_t st, int* a, int n);
extern "C" hipError_t fc_launch_zero_f4(hipStream_t st, float4* a, int n);
extern "C" int fc_lds_bytes(int nmax, int mmax, int wmax, int block, int learned, int h2lds);
extern "C" hipError_t fc_launch_split_step(hipStream_t st, const fcd::BatchDev* B, const fcd::RunParams* P,
                                           const int* node_sys, int ntot_nodes, int maxN, int do_move, int step);
extern "C" hipError_t fc_launch_reset(hipStream_t st, const fcd::BatchDev* B, const fcd::RunParams* P);
extern "C" hipError_t fc_launch_precond(hipStream_t st, const fcd::BatchDev* B, const fcd::RunParams* P, float4* out,
                                        const float* gnode, const float* gcomp);
extern "C" hipError_t fc_persistent_occupancy(int block, int hb, int pc, size_t lds, int* nblocks);
extern "C" hipError_t fc_launch_reset_subset(hipStream_t st, const fcd::BatchDev* B, const fcd::RunParams* P,
                                             const int* only);

namespace fcw {

bool hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    std::fprintf(stderr, "[libMMFFmulti_lib] HIP error in %s: %s\n", what, hipGetErrorString(e));
    return false;
}

template <class T>
void DevArr<T>::alloc(size_t cnt) {
    if (cnt == n && p) return;
    free_();
    n = cnt;
    if (cnt == 0) return;
    hip_ok(hipMalloc((void**)&p, cnt * sizeof(T)), "hipMalloc");
}
template <class T>
void DevArr<T>::free_() {
    if (p) hipFree(p);
    p = nullptr;
    n = 0;
}
template <class T>
void DevArr<T>::upload(const std::vector<T>& h, hipStream_t s) {
    alloc(std::max<size_t>(h.size(), 1));
    if (!h.empty()) hip_ok(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
}
template <class T>
void DevArr<T>::download(std::vector<T>& h, hipStream_t s) const {
    size_t cnt = std::min(h.size(), n);
    if (cnt) hip_ok(hipMemcpyAsync(h.data(), p, cnt * sizeof(T), hipMemcpyDeviceToHost, s), "download");
}

// explicit instantiations (the other translation units use these)
template struct DevArr<float4>;
template struct DevArr<int4>;
template struct DevArr<int>;
template struct DevArr<unsigned>;
template struct DevArr<unsigned short>;
template struct DevArr<float>;
template struct DevArr<double>;
template struct DevArr<long long>;
template struct DevArr<fcd::SysInfo>;
template struct DevArr<fcd::PairCorrDev>;

World& world() {
    static World W;
    return W;
}

void World::clear() {
    mols.clear();
    info.clear();
    nSys = ntotA = ntotM = ntotV = maxN = maxM = 0;
}

static inline float4 f4(double a, double b, double c, double d) {
    return make_float4((float)a, (float)b, (float)c, (float)d);
}

bool World::set_systems(std::vector<fc::MolSys>&& ms) {
    mols = std::move(ms);
    nSys = (int)mols.size();
    if (nSys == 0) return false;
    ffl = mols[0];
    build_packed();
    register_buffers();
    return true;
}

bool World::set_surface_atoms(const fc::Surface& s, const int* nPBC) {
    surf_pos.clear();
    surf_REQ.clear();
    surf_shifts.clear();
    surf_host = s;
    for (int k = 0; k < 3; k++) surf_nPBC[k] = nPBC[k];
    const int n = s.natoms();
    for (int i = 0; i < n; i++) {
        surf_pos.push_back(f4(s.apos[i * 3], s.apos[i * 3 + 1], s.apos[i * 3 + 2], 0.0));
        surf_REQ.push_back(f4(s.REQs[i * 4], s.REQs[i * 4 + 1], s.REQs[i * 4 + 2], s.REQs[i * 4 + 3]));
    }
    if (n > 0) {  // GridFF::makePBCshifts(nPBC, lvec) (NBFF.h:1142-1156)
        std::vector<double> sh;
        fc::make_pbc_shifts(s.lvec, nPBC, sh);
        for (size_t k = 0; k < sh.size() / 3; k++) surf_shifts.push_back(f4(sh[k * 3], sh[k * 3 + 1], sh[k * 3 + 2], 0.0));
        bGridFF = false;  // the substrate enters through its atoms
    }
    if (!ensure_device()) return false;
    d_surf_pos.upload(surf_pos, stream);
    d_surf_REQ.upload(surf_REQ, stream);
    d_surf_shifts.upload(surf_shifts, stream);
    if (nSys > 0) plan_classes();
    return hip_ok(hipStreamSynchronize(stream), "surface atoms");
}

bool World::set_lattice(int isys, const double* lvec, const int* nPBC) {
    if (nSys == 0 || isys >= nSys) return false;
    if (dev_ready) {  // keep the current state of every system
        if (!download_all(true, true)) return false;
        for (int s = 0; s < nSys; s++) unpack_system(s, true, true);
    }
    for (int s = (isys < 0 ? 0 : isys); s < (isys < 0 ? nSys : isys + 1); s++) {
        fc::MolSys& m = mols[s];
        m.pbc = lvec != nullptr;
        if (lvec)
            for (int k = 0; k < 9; k++) m.lvec[k] = lvec[k];
        if (nPBC)
            for (int k = 0; k < 3; k++) m.nPBC[k] = std::max(nPBC[k], 0);
    }
    const std::vector<float4> keep_constr = constr, keep_constrK = constrK, keep_bboxes = bboxes;
    build_packed();
    constr = keep_constr;
    constrK = keep_constrK;
    bboxes = keep_bboxes;
    register_buffers();
    return dev_ready ? upload_all(true, true, true) : true;
}

void World::rebuild_exclusions() {
    excl.clear();
    excl_rot.clear();
    corr.clear();
    ring_ok = true;
    const bool sub13 = bNonBonded && sw.doAngles;
    for (int s = 0; s < nSys; s++) {
        const fc::MolSys& m = mols[s];
        std::vector<uint32_t> mask;
        std::vector<fc::PairCorr> pc;
        int W = 0;
        fc::make_exclusions(m.natoms, m.nnode, m.neighs.data(), sub13, mask, W, pc);
        info[s].mask0 = (int)excl.size();
        info[s].corr0 = (int)corr.size();
        info[s].ncorr = (int)pc.size();
        excl.insert(excl.end(), mask.begin(), mask.end());
        std::vector<uint32_t> rot;
        if (!fc::rotate_exclusions(m.natoms, mask, W, rot)) ring_ok = false;
        excl_rot.insert(excl_rot.end(), rot.begin(), rot.end());
        if (!m.pbc) {
            for (auto& c : pc) corr.push_back({c.i, c.j, c.w, -1});
            continue;
        }
        // Periodic molecule: the image loop skips a bonded neighbour only at the
        // image of the bond (NBFF::evalLJQs_ng4_PBC_atom_omp, NBFF.h:575-620), and
        // every 1-3 pair is subtracted once per angle path at the image reached
        // through the bonds (MMFFsp3_loc::eval_atom, MMFFsp3_loc.h:525-536), which
        // may be an image outside the summed range or the image of a 1-2 bond of
        // the same two atoms in a short cell.  One directed entry per path and
        // end atom (w = -1, halved with the rest of the pair energy); `path`
        // names the node and the two slots, whose neighbour images give the shift.
        info[s].ncorr = 0;
        if (!sub13) continue;
        for (int c = 0; c < m.nnode; c++) {
            const int* ng = m.neighs.data() + c * 4;
            for (int a = 0; a < 3; a++) {
                if (ng[a] < 0) break;
                for (int b = a + 1; b < 4; b++) {
                    if (ng[b] < 0) break;
                    corr.push_back({ng[a], ng[b], -1.f, c * 16 + a * 4 + b});
                    corr.push_back({ng[b], ng[a], -1.f, c * 16 + b * 4 + a});
                    info[s].ncorr += 2;
                }
            }
        }
    }
}

void World::build_packed() {
    info.assign(nSys, fcd::SysInfo{});
    ntotA = ntotM = ntotV = maxN = maxM = 0;
    bool uniform = true;
    for (int s = 0; s < nSys; s++) {
        const fc::MolSys& m = mols[s];
        info[s].natoms = m.natoms;
        info[s].nnode = m.nnode;
        info[s].a0 = ntotA;
        info[s].n0 = ntotM;
        info[s].v0 = ntotV;
        ntotA += m.natoms;
        ntotM += m.nnode;
        ntotV += m.natoms + m.nnode;
        maxN = std::max(maxN, m.natoms);
        maxM = std::max(maxM, m.nnode);
        if (m.natoms != mols[0].natoms || m.nnode != mols[0].nnode) uniform = false;
    }
    natoms_uniform = uniform ? mols[0].natoms : 0;
    nnode_uniform = uniform ? mols[0].nnode : 0;
    nvecs_uniform = uniform ? mols[0].nvecs() : 0;
    atoms.assign(ntotV, f4(0, 0, 0, 0));
    aforces.assign(ntotV, f4(0, 0, 0, 0));
    avel.assign(ntotV, f4(0, 0, 0, 0));
    REQs.assign(ntotA, f4(0, 0, 0, 0));
    PLQs.assign(ntotA, f4(0, 0, 0, 0));
    constr.assign(ntotA, f4(-1, -1, -1, -1));
    constrK.assign(ntotA, f4(-1, -1, -1, -1));
    bboxes.assign((size_t)nSys * 3, f4(0, 0, 0, 0));
    for (int s = 0; s < nSys; s++) {
        bboxes[s * 3] = f4(-1e8, -1e8, -1e8, 0);
        bboxes[s * 3 + 1] = f4(1e8, 1e8, 1e8, 0);
    }
    sysflags.assign(nSys, 0);
    neighs.assign(ntotA, make_int4(-1, -1, -1, -1));
    neighCell.assign(ntotA, make_int4(-1, -1, -1, -1));
    bkNeighs.assign(ntotV, make_int4(-1, -1, -1, -1));
    MMpars.assign(ntotM, f4(0, 0, 0, 0));
    BLs = BKs = Ksp = Kpp = MMpars;
    node_sys.assign(ntotM, 0);
    elem.assign(ntotA, 0);
    comp.assign(ntotA, 0);
    ncomp.assign(nSys, 1);
    lvecs.assign((size_t)nSys * 3, f4(0, 0, 0, 0));
    ilvecs.assign((size_t)nSys * 3, f4(0, 0, 0, 0));
    // image shifts: npbc per periodic system, one zero shift per vacuum system
    // (the reference's pbcshifts layout, nSys*npbc, for uniform batches)
    int nsh = 0;
    n_periodic = 0;
    for (int s = 0; s < nSys; s++) {
        const fc::MolSys& m = mols[s];
        info[s].sh0 = nsh;
        info[s].npbc = m.pbc ? (2 * m.nPBC[0] + 1) * (2 * m.nPBC[1] + 1) * (2 * m.nPBC[2] + 1) : 0;
        nsh += std::max(info[s].npbc, 1);
        n_periodic += m.pbc ? 1 : 0;
    }
    pbcshifts.assign((size_t)nsh, f4(0, 0, 0, 0));
    nsteps.assign(nSys, 0);
    status.assign(nSys, 0);
    Esys.assign(nSys, 0.0);
    Fmax.assign(nSys, 0.f);
    for (int s = 0; s < nSys; s++) pack_system(s, true, false, false, true);
    // The reference's Python view maps gpu_neighs / gpu_neighCell as
    // (nSys, nvecs, 4) (pyBall/MMFF_multi.py:233-235) although they hold
    // natoms rows per system: for uniform batches the host mirrors are padded
    // with -1 rows to nSys*nvecs so that such a view stays inside the buffer.
    if (nvecs_uniform > 0) {
        const size_t rows = std::max((size_t)ntotA, (size_t)nSys * nvecs_uniform);
        neighs.resize(rows, make_int4(-1, -1, -1, -1));
        neighCell.resize(rows, make_int4(-1, -1, -1, -1));
    }
    rebuild_exclusions();
    // template views
    ndims[0] = ffl.nvecs() * 3;
    ndims[1] = ffl.nnode;
    ndims[2] = ffl.natoms - ffl.nnode;
    ndims[3] = ffl.nvecs();
    ndims[4] = 0;
    ndims[5] = 0;
    DOFs = ffl.apos;
    fDOFs.assign(DOFs.size(), 0.0);
    vDOFs.assign(DOFs.size(), 0.0);
    tmpl_neighs = ffl.neighs;
    selection.assign(std::max(ffl.natoms, 1), 0);
}

// pack one system (double, MMFFsp3_loc layout) into the float mirrors
void World::pack_system(int s, bool bParams, bool bForces, bool bVel, bool blvec) {
    const fc::MolSys& m = mols[s];
    const fcd::SysInfo& I = info[s];
    const int N = m.natoms, M = m.nnode, V = N + M;
    for (int v = 0; v < V; v++) atoms[I.v0 + v] = f4(m.apos[v * 3], m.apos[v * 3 + 1], m.apos[v * 3 + 2], 0.0);
    if (bForces && (int)m.fapos.size() == V * 3)
        for (int v = 0; v < V; v++) aforces[I.v0 + v] = f4(m.fapos[v * 3], m.fapos[v * 3 + 1], m.fapos[v * 3 + 2], 0.0);
    if (bVel && (int)m.vapos.size() == V * 3)
        for (int v = 0; v < V; v++) avel[I.v0 + v] = f4(m.vapos[v * 3], m.vapos[v * 3 + 1], m.vapos[v * 3 + 2], 0.0);
    std::vector<double> shifts;
    if (m.pbc) fc::make_pbc_shifts(m.lvec, m.nPBC, shifts);
    if (blvec) {
        // vacuum molecules: zero lattice, one zero shift; periodic ones: lattice
        // rows, the transposed inverse (Mat3_to_cl(ff.invLvec), setLvec's
        // invert_T_to, MMFFsp3_loc.h:232) and the image shifts
        for (int k = 0; k < 3; k++) {
            lvecs[s * 3 + k] = f4(0, 0, 0, 0);
            ilvecs[s * 3 + k] = f4(0, 0, 0, 0);
        }
        pbcshifts[I.sh0] = f4(0, 0, 0, 0);
        if (m.pbc) {
            const double* L = m.lvec;
            double inv[9];
            const double det = L[0] * (L[4] * L[8] - L[5] * L[7]) - L[1] * (L[3] * L[8] - L[5] * L[6]) +
                               L[2] * (L[3] * L[7] - L[4] * L[6]);
            const double id = det != 0.0 ? 1.0 / det : 0.0;
            inv[0] = (L[4] * L[8] - L[5] * L[7]) * id;
            inv[1] = (L[2] * L[7] - L[1] * L[8]) * id;
            inv[2] = (L[1] * L[5] - L[2] * L[4]) * id;
            inv[3] = (L[5] * L[6] - L[3] * L[8]) * id;
            inv[4] = (L[0] * L[8] - L[2] * L[6]) * id;
            inv[5] = (L[2] * L[3] - L[0] * L[5]) * id;
            inv[6] = (L[3] * L[7] - L[4] * L[6]) * id;
            inv[7] = (L[1] * L[6] - L[0] * L[7]) * id;
            inv[8] = (L[0] * L[4] - L[1] * L[3]) * id;
            for (int k = 0; k < 3; k++) {
                lvecs[s * 3 + k] = f4(L[k * 3], L[k * 3 + 1], L[k * 3 + 2], 0);
                ilvecs[s * 3 + k] = f4(inv[k], inv[3 + k], inv[6 + k], 0);
            }
            for (int k = 0; k < I.npbc; k++)
                pbcshifts[I.sh0 + k] = f4(shifts[k * 3], shifts[k * 3 + 1], shifts[k * 3 + 2], 0);
        }
    }
    if (!bParams) return;
    const double alpha = grid.alphaMorse;
    for (int i = 0; i < N; i++) {
        const double* q = &m.REQs[i * 4];
        REQs[I.a0 + i] = f4(q[0], q[1], q[2], q[3]);
        // REQ2PLQ_d (Forces.h:611-619) in double, rounded once to float
        const double e = std::exp(alpha * q[0]);
        const double cL = e * q[1];
        PLQs[I.a0 + i] = f4(e * cL, cL, q[2], e * e * q[3]);
        neighs[I.a0 + i] = make_int4(m.neighs[i * 4], m.neighs[i * 4 + 1], m.neighs[i * 4 + 2], m.neighs[i * 4 + 3]);
        neighCell[I.a0 + i] = make_int4(-1, -1, -1, -1);
        elem[I.a0 + i] = m.elem.empty() ? 0 : m.elem[i];
    }
    if (m.pbc) {  // image of each bonded neighbour at the current geometry (makeNeighCells)
        std::vector<int> nc((size_t)N * 4);
        fc::make_neigh_cells(N, m.apos.data(), m.neighs.data(), I.npbc, shifts.data(), nc.data());
        for (int i = 0; i < N; i++)
            neighCell[I.a0 + i] = make_int4(nc[i * 4], nc[i * 4 + 1], nc[i * 4 + 2], nc[i * 4 + 3]);
    }
    for (int k = 0; k < M; k++) {
        auto P4 = [&](const std::vector<double>& a) { return f4(a[k * 4], a[k * 4 + 1], a[k * 4 + 2], a[k * 4 + 3]); };
        MMpars[I.n0 + k] = P4(m.apars);
        BLs[I.n0 + k] = P4(m.bLs);
        BKs[I.n0 + k] = P4(m.bKs);
        Ksp[I.n0 + k] = P4(m.Ksp);
        Kpp[I.n0 + k] = P4(m.Kpp);
        node_sys[I.n0 + k] = s;
    }
    ncomp[s] = fc::connected_components(N, m.neighs.data(), fcd::PC_MAXC, comp.data() + I.a0);
    std::vector<int> bk((size_t)N * 4);
    if (!fc::make_back_neighs(N, M, m.neighs.data(), bk.data()))
        std::fprintf(stderr, "[libMMFFmulti_lib] system %d: an atom has more than 4 back-neighbours\n", s);
    const int base = I.n0 * 8;
    auto add = [&](int x, int off) { return x < 0 ? -1 : x + off; };
    for (int i = 0; i < N; i++)
        bkNeighs[I.v0 + i] = make_int4(add(bk[i * 4], base), add(bk[i * 4 + 1], base), add(bk[i * 4 + 2], base),
                                       add(bk[i * 4 + 3], base));
    for (int k = 0; k < M; k++)
        bkNeighs[I.v0 + N + k] = make_int4(add(bk[k * 4], base + 4 * M), add(bk[k * 4 + 1], base + 4 * M),
                                           add(bk[k * 4 + 2], base + 4 * M), add(bk[k * 4 + 3], base + 4 * M));
}

void World::unpack_system(int s, bool bForces, bool bVel) {
    fc::MolSys& m = mols[s];
    const fcd::SysInfo& I = info[s];
    const int V = m.natoms + m.nnode;
    for (int v = 0; v < V; v++) {
        m.apos[v * 3] = atoms[I.v0 + v].x;
        m.apos[v * 3 + 1] = atoms[I.v0 + v].y;
        m.apos[v * 3 + 2] = atoms[I.v0 + v].z;
    }
    if (bForces) {
        m.fapos.resize((size_t)V * 3);
        for (int v = 0; v < V; v++) {
            m.fapos[v * 3] = aforces[I.v0 + v].x;
            m.fapos[v * 3 + 1] = aforces[I.v0 + v].y;
            m.fapos[v * 3 + 2] = aforces[I.v0 + v].z;
        }
    }
    if (bVel) {
        m.vapos.resize((size_t)V * 3);
        for (int v = 0; v < V; v++) {
            m.vapos[v * 3] = avel[I.v0 + v].x;
            m.vapos[v * 3 + 1] = avel[I.v0 + v].y;
            m.vapos[v * 3 + 2] = avel[I.v0 + v].z;
        }
    }
}

bool World::ensure_device() {
    if (dev_ready) return true;
    int ndev = 0;
    if (!hip_ok(hipGetDeviceCount(&ndev), "hipGetDeviceCount") || ndev < 1) {
        std::fprintf(stderr, "[libMMFFmulti_lib] no HIP device available: this library has no CPU path\n");
        return false;
    }
    if (!hip_ok(hipStreamCreate(&stream), "hipStreamCreate")) return false;
    for (int i = 0; i < 4; i++) {
        hipStream_t st;
        if (!hip_ok(hipStreamCreate(&st), "hipStreamCreate")) return false;
        class_streams.push_back(st);
    }
    // allocation-free relaxation: this device's scratch array and the fork/join
    // events of its class launches, once at set-up
    if (!hip_ok(fc_noalloc_scratch(&noalloc_ptr), "scratch address")) return false;
    for (int i = 0; i < 10; i++) {
        hipEvent_t e = nullptr;
        if (!hip_ok(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate")) return false;
        noalloc_events.push_back(e);
    }
    dev_ready = true;
    return true;
}

// Per-system flags of the restraint terms: bit 0 = a restrained atom
// (constr.w > 0), bit 1 = a wall (positive stiffness on some axis).
void World::update_sysflags() {
    sysflags.assign(nSys, 0);
    for (int s = 0; s < nSys; s++) {
        const fcd::SysInfo& I = info[s];
        for (int i = 0; i < I.natoms; i++)
            if (constr[I.a0 + i].w > 0.f) {
                sysflags[s] |= 1;
                break;
            }
        const float4 k = bboxes[s * 3 + 2];
        if (k.x > 0.f || k.y > 0.f || k.z > 0.f) sysflags[s] |= 2;
    }
}

bool World::upload_all(bool bParams, bool bForces, bool bVel) {
    if (!ensure_device()) return false;
    update_sysflags();
    d_constr.upload(constr, stream);
    d_constrK.upload(constrK, stream);
    d_bboxes.upload(bboxes, stream);
    d_sysflags.upload(sysflags, stream);
    d_atoms.upload(atoms, stream);
    d_atoms0.upload(atoms, stream);
    if (bForces || !d_aforce.p) d_aforce.upload(aforces, stream);
    if (bVel || !d_avel.p) d_avel.upload(avel, stream);
    if (bParams || !d_REQs.p) {
        d_REQs.upload(REQs, stream);
        d_PLQs.upload(PLQs, stream);
        d_neighs.upload(neighs, stream);
        d_bkNeighs.upload(bkNeighs, stream);
        d_neighCell.upload(neighCell, stream);
        d_pbcshifts.upload(pbcshifts, stream);
        d_apars.upload(MMpars, stream);
        d_bLs.upload(BLs, stream);
        d_bKs.upload(BKs, stream);
        d_Ksp.upload(Ksp, stream);
        d_Kpp.upload(Kpp, stream);
        d_excl.upload(excl, stream);
        d_excl_rot.upload(excl_rot, stream);
        d_corr.upload(corr, stream);
        d_info.upload(info, stream);
        d_node_sys.upload(node_sys, stream);
        d_comp.upload(comp, stream);
        d_gridcache.alloc(std::max<size_t>((size_t)ntotA * 16, 1));
        d_ncomp.upload(ncomp, stream);
        d_elem.upload(elem, stream);
        d_fneigh.alloc(std::max<size_t>((size_t)ntotM * 8, 1));
        d_ebond.alloc(std::max<size_t>((size_t)ntotM, 1));
        d_enb.alloc(std::max<size_t>((size_t)ntotA, 1));
        d_fire.alloc(std::max(nSys, 1));
        d_stall.alloc(std::max(nSys, 1));
        d_nsteps.upload(nsteps, stream);
        d_status.upload(status, stream);
        d_Esys.upload(Esys, stream);
        d_Fmax.upload(Fmax, stream);
        d_queue.alloc(8);
        plan_classes();
    }
    if (bGridFF && !grid_f4.empty() && d_grid.n != grid_f4.size()) d_grid.upload(grid_f4, stream);
    return hip_ok(hipStreamSynchronize(stream), "upload");
}

bool World::download_all(bool bForces, bool bVel) {
    if (!dev_ready) return false;
    d_atoms.download(atoms, stream);
    if (bForces) d_aforce.download(aforces, stream);
    if (bVel) d_avel.download(avel, stream);
    d_nsteps.download(nsteps, stream);
    d_status.download(status, stream);
    d_Esys.download(Esys, stream);
    d_Fmax.download(Fmax, stream);
    bool ok = hip_ok(hipStreamSynchronize(stream), "download");
    // refresh the template views from system 0
    if (ok && nSys > 0) {
        const fcd::SysInfo& I = info[0];
        const int V = I.natoms + I.nnode;
        for (int v = 0; v < V; v++) {
            DOFs[v * 3] = atoms[I.v0 + v].x;
            DOFs[v * 3 + 1] = atoms[I.v0 + v].y;
            DOFs[v * 3 + 2] = atoms[I.v0 + v].z;
            fDOFs[v * 3] = aforces[I.v0 + v].x;
            fDOFs[v * 3 + 1] = aforces[I.v0 + v].y;
            fDOFs[v * 3 + 2] = aforces[I.v0 + v].z;
            vDOFs[v * 3] = avel[I.v0 + v].x;
            vDOFs[v * 3 + 1] = avel[I.v0 + v].y;
            vDOFs[v * 3 + 2] = avel[I.v0 + v].z;
        }
        Es[0] = Esys[0];
    }
    return ok;
}

bool World::set_grid(const fc::GridFFHost& g) {
    grid = g;
    const size_t np = (size_t)g.nx * g.ny * g.nz;
    grid_f4.resize(np);
    // Bspline_to_f4 (GridFF.h:1219-1227)
    for (size_t i = 0; i < np; i++)
        grid_f4[i] = f4(g.coefs[i * 3], g.coefs[i * 3 + 1], g.coefs[i * 3 + 2], 0.0);
    bGridFF = true;
    // PLQs depend on alphaMorse
    for (int s = 0; s < nSys; s++) pack_system(s, true, false, false, false);
    if (dev_ready) {
        d_grid.upload(grid_f4, stream);
        d_PLQs.upload(PLQs, stream);
        return hip_ok(hipStreamSynchronize(stream), "set_grid");
    }
    return true;
}

fcd::BatchDev World::batch_dev() {
    fcd::BatchDev B{};
    B.nsys = nSys;
    B.info = d_info.p;
    B.atoms = d_atoms.p;
    B.avel = d_avel.p;
    B.aforce = d_aforce.p;
    B.REQs = d_REQs.p;
    B.PLQs = d_PLQs.p;
    B.neighs = d_neighs.p;
    B.bkNeighs = d_bkNeighs.p;
    B.apars = d_apars.p;
    B.bLs = d_bLs.p;
    B.bKs = d_bKs.p;
    B.Ksp = d_Ksp.p;
    B.Kpp = d_Kpp.p;
    B.excl = d_excl.p;
    B.excl_rot = d_excl_rot.p;
    B.corr = d_corr.p;
    B.fneigh = d_fneigh.p;
    B.ebond = d_ebond.p;
    B.enb = d_enb.p;
    B.nsteps = d_nsteps.p;
    B.status = d_status.p;
    B.Esys = d_Esys.p;
    B.Fmax = d_Fmax.p;
    B.fire = d_fire.p;
    B.stall = d_stall.p;
    B.comp = d_comp.p;
    B.ncomp = d_ncomp.p;
    B.elem = d_elem.p;
    B.gnode = d_gnode.p;
    B.gcomp = d_gcomp.p;
    B.systime = (record_times && d_systime.p) ? d_systime.p : nullptr;
    B.constr = d_constr.p;
    B.constrK = d_constrK.p;
    B.bbox = d_bboxes.p;
    B.sysflags = d_sysflags.p;
    B.gridcache = d_gridcache.p;
    B.gridcache_stride = ntotA;
    B.pbcshifts = d_pbcshifts.p;
    B.neighCell = d_neighCell.p;
    B.snap = (snap_on && d_snap.p) ? d_snap.p : nullptr;
    B.snap_every = std::max(snap_every, 1);
    B.nsnap = nsnap;
    B.ntotV = ntotV;
    return B;
}

fcd::RunParams World::run_params(double Fconv, int nstep_max) {
    fcd::RunParams P{};
    P.sw = sw;
    if (!bMMFF) P.sw = fcd::FFSwitch{0, 0, 0, 0};
    P.bNonBonded = bNonBonded ? 1 : 0;
    P.bGridFF = (bGridFF && d_grid.p) ? 1 : 0;
    P.grid.coefs = d_grid.p;
    P.grid.nx = grid.nx;
    P.grid.ny = grid.ny;
    P.grid.nz = grid.nz;
    P.grid.p0x = (float)grid.pos0[0];
    P.grid.p0y = (float)grid.pos0[1];
    P.grid.p0z = (float)grid.pos0[2];
    P.grid.isx = (float)(1.0 / grid.step[0]);
    P.grid.isy = (float)(1.0 / grid.step[1]);
    P.grid.isz = (float)(1.0 / grid.step[2]);
    P.fire = fp;
    P.Fconv2 = (float)(Fconv * Fconv);
    P.conv_mode = conv_mode;
    P.Flimit = Flimit;
    P.nstep_max = nstep_max;
    P.pc_mode = pc_mode;
    P.pc_gstar = pc_gstar;
    P.pc_grigid = pc_grigid;
    P.pc_every = std::max(pc_every, 1);
    P.pc_stall = std::max(pc_stall, 0);
    if (!surf_pos.empty()) {
        P.surf_pos = d_surf_pos.p;
        P.surf_REQ = d_surf_REQ.p;
        P.surf_shifts = d_surf_shifts.p;
        P.nsurf = (int)surf_pos.size();
        P.nsurf_pbc = (int)surf_shifts.size();
        P.surf_K = (float)-surf_alpha;
    }
    if (d_mlpW.p) {
        P.mlp.W1 = d_mlpW.p;
        P.mlp.W2 = d_mlpW.p + 64 * 16;
        P.mlp.W3 = d_mlpW.p + 64 * 16 + 64 * 128;
        P.mlp.b1 = d_mlpF.p;
        P.mlp.b2 = d_mlpF.p + 64;
        P.mlp.b3 = d_mlpF.p + 128;
        P.mlp.W4 = d_mlpF.p + 192;
        P.mlp.b4 = d_mlpF.p + 192 + 128;
    }
    return P;
}

// Group systems by workgroup width (64/128/192/256 threads: the smallest
// that holds one atom and one node per thread and two vecs per thread) and
// order each group by decreasing cost (largest first).  The device's resident
// workgroup slots are split between the classes in proportion to their
// estimated time; workgroups that run out of work steal from smaller classes.
// The estimate uses the previous relaxation's step counts of the same batch
// when available (the autotuning loop), else a size-based guess.
void World::plan_classes(bool upload) {
    classes.clear();
    const int blocks[4] = {64, 128, 192, 256};
    std::vector<SizeClass> cl(4);
    for (int c = 0; c < 4; c++) cl[c].block = blocks[c];
    hb_batch = false;
    for (const float4& q : REQs)
        if (q.w < 0.f) hb_batch = true;
    int too_big = 0;
    sys_class.assign(nSys, -1);
    for (int s = 0; s < nSys; s++) {
        const int N = info[s].natoms, M = info[s].nnode;
        int c = -1;
        for (int k = 0; k < 4; k++)
            if (((class_mask >> k) & 1 || k == 3) && N <= blocks[k] && M <= blocks[k] && N + M <= 2 * blocks[k]) {
                c = k;
                break;
            }
        // periodic molecules and explicit substrate atoms: the split kernels
        if (mols[s].pbc || !surf_pos.empty()) c = -1;
        sys_class[s] = c;
        if (c < 0) {
            too_big++;
            continue;
        }
        cl[c].order.push_back(s);
        cl[c].nmax = std::max(cl[c].nmax, N);
        cl[c].mmax = std::max(cl[c].mmax, M);
    }
    n_unclassed = too_big;
    if (too_big && verbosity > 0)
        std::fprintf(stderr, "[libMMFFmulti_lib] %d systems (periodic, over substrate atoms, or above the 256-thread "
                             "workgroup) are relaxed by the split kernels (plain FIRE)\n", too_big);
    int ncu = 256;
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) == hipSuccess) ncu = prop.multiProcessorCount;
    }
    // LDS of a class covers every system it may steal
    int nrun = 0, mrun = 0;
    std::vector<int> qoff(4, 0), qcnt(4, 0);
    std::vector<int> allorder;
    for (int c = 0; c < 4; c++) {
        SizeClass& C = cl[c];
        nrun = std::max(nrun, C.nmax);
        mrun = std::max(mrun, C.mmax);
        C.nmax = nrun;
        C.mmax = mrun;
        C.wmax = ((C.nmax >> 1) >> 5) + 1;  // exclusion words of the offsets k <= nmax/2
        // learned gains: the second activation buffer goes to LDS when the class
        // keeps two resident workgroups per CU pair of SIMDs (see plan below)
        C.h2lds = 0;
        C.lds = (size_t)fc_lds_bytes(std::max(C.nmax, 1), std::max(C.mmax, 1), std::max(C.wmax, 1), C.block,
                                     pc_mode == 2 ? 1 : 0, 0);
        if (pc_mode == 2 && h2lds_allowed) {
            const size_t budget = (size_t)160 * 1024 / std::max(1, 8 / (C.block / 64));
            const size_t with = (size_t)fc_lds_bytes(std::max(C.nmax, 1), std::max(C.mmax, 1), std::max(C.wmax, 1),
                                                     C.block, 1, 1);
            if (with <= budget) {
                C.h2lds = 1;
                C.lds = with;
            }
        }
        auto cost = [&](int s) {
            const double N = info[s].natoms;
            const double st = (cost_feedback && nsteps.size() == (size_t)nSys && nsteps[s] > 0)
                                  ? nsteps[s] + 1.0
                                  : std::max(1.0, cost_c0 + cost_c1 * N);
            return st * N * N;
        };
        std::sort(C.order.begin(), C.order.end(), [&](int a, int b) {
            double ca = cost(a), cb = cost(b);
            return ca != cb ? ca > cb : a < b;
        });
        // time estimate: work over the busy fraction of the workgroup's lanes
        C.work = 0;
        for (int s : C.order) C.work += cost(s) * C.block / std::max(1, info[s].natoms);
        qoff[c] = (int)allorder.size();
        qcnt[c] = (int)C.order.size();
        allorder.insert(allorder.end(), C.order.begin(), C.order.end());
    }
    double wtot = 0;
    for (auto& C : cl) wtot += C.work;
    for (int c = 0; c < 4; c++) {
        SizeClass& C = cl[c];
        C.cls = c;
        if (C.order.empty()) continue;
        int nb = 1;
        fc_persistent_occupancy(C.block, hb_batch ? 1 : 0, (pc_mode == 2 && !learned_inkernel) ? 3 : pc_mode, C.lds, &nb);
        nb = std::max(nb, 1);
        C.occ = nb;
        int g = (int)std::lround((double)ncu * nb * C.work / std::max(wtot, 1e-30));
        if (sched_mode == 1) g = ncu * nb;  // every class may fill the device; the dispatcher interleaves them
        // twice the resident width: the extra workgroups are dispatched as
        // other classes retire and drain the remaining queue
        if (sched_mode == 2) g = 2 * ncu * nb;
        C.ngroups = std::max(1, std::min(g, (int)C.order.size()));
        classes.push_back(C);
    }
    // launch the most expensive class first
    std::stable_sort(classes.begin(), classes.end(), [](const SizeClass& a, const SizeClass& b) { return a.work > b.work; });
    qinfo.assign(8, 0);
    for (int c = 0; c < 4; c++) {
        qinfo[c] = qoff[c];
        qinfo[4 + c] = qcnt[c];
    }
    if (allorder.empty()) allorder.push_back(0);
    h_order = allorder;
    if (upload) d_order.upload(allorder, stream);
}

int World::relax_persistent(int nstep_max, double Fconv, bool restore_start) {
    last_chunks = 0;
    if (!dev_ready || nSys == 0) return -1;
    if (pc_mode == 2 && !learned_inkernel) return relax_learned(nstep_max, Fconv, restore_start);
    if (pc_mode == 2 && !d_mlpW.p) {
        std::fprintf(stderr, "[libMMFFmulti_lib] learned gains need proposal weights\n");
        return -1;
    }
    if (!ring_ok) {
        std::fprintf(stderr, "[libMMFFmulti_lib] asymmetric exclusions: relaxing with the split kernels\n");
        return relax_split(nstep_max, Fconv, restore_start);
    }
    if (autotune) plan_classes();
    if (restore_start)
        hip_ok(hipMemcpyAsync(d_atoms.p, d_atoms0.p, sizeof(float4) * ntotV, hipMemcpyDeviceToDevice, stream),
               "restore");
    if (record_times) d_systime.alloc((size_t)2 * nSys);
    fc_launch_zero_f4(stream, d_stall.p, nSys);  // fallback flags of this relaxation only
    fcd::BatchDev B = batch_dev();
    fcd::RunParams P = run_params(Fconv, nstep_max);
    std::vector<int> zeros(8, 0);
    hip_ok(hipMemcpyAsync(d_queue.p, zeros.data(), sizeof(int) * 8, hipMemcpyHostToDevice, stream), "queue reset");
    // learned gains evaluated in-kernel: one global activation slot per workgroup
    std::vector<size_t> slot_off(classes.size(), 0);
    if (pc_mode == 2) {
        size_t tot = 0;
        for (size_t k = 0; k < classes.size(); k++) {
            slot_off[k] = tot;
            tot += (size_t)classes[k].ngroups * fc_hslot_elems(std::max(classes[k].nmax, 1));
        }
        d_hslot.alloc(std::max<size_t>(tot, 1));
    }
    EventSet events;
    hipEvent_t e0 = events.make(), e1 = events.make();
    hipEventRecord(e0, stream);
    std::vector<hipEvent_t> starts, done;
    for (size_t k = 0; k < classes.size(); k++) {
        SizeClass& C = classes[k];
        hipStream_t st = class_streams[k % class_streams.size()];
        hipStreamWaitEvent(st, e0, 0);
        hipEvent_t es = events.make(), ev = events.make();
        hipEventRecord(es, st);
        hipError_t err = fc_launch_relax_persistent(C.block, hb_batch ? 1 : 0, C.ngroups, C.lds, st, &B, &P, d_order.p,
                                                    qinfo.data(), d_queue.p, C.cls, C.nmax, C.mmax, C.wmax,
                                                    sched_mode == 0 ? 1 : 0,
                                                    pc_mode == 2 ? d_hslot.p + slot_off[k] : nullptr, C.h2lds);
        if (!hip_ok(err, "relax_persistent launch")) return -1;
        hipEventRecord(ev, st);
        hipStreamWaitEvent(stream, ev, 0);
        starts.push_back(es);
        done.push_back(ev);
    }
    hipEventRecord(e1, stream);
    if (!hip_ok(hipStreamSynchronize(stream), "relax_persistent")) return -1;
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    last_ms = ms;
    for (size_t c = 0; c < classes.size(); c++) {
        float cm = 0;
        hipEventElapsedTime(&cm, starts[c], done[c]);
        classes[c].last_ms = cm;
    }
    if (n_unclassed > 0 && !relax_unclassed(nstep_max, Fconv)) return -1;
    d_nsteps.download(nsteps, stream);
    d_status.download(status, stream);
    if (!hip_ok(hipStreamSynchronize(stream), "relax results")) return -1;
    int mx = 0;
    for (int s = 0; s < nSys; s++) mx = std::max(mx, nsteps[s]);
    return mx;
}

// relax_persistent without allocation: the queue counters, the queue order, the
// in-kernel network's activation slots and the GridFF window cache of this
// relaxation live in the library's static scratch array (noalloc_ptr, looked up
// at set-up); the class launches fork and join through events made at set-up.
// Calls no hipMalloc/hipFree.  Paths that need more (systems above the largest
// workgroup class, asymmetric exclusions, the chunked learned form, per-system
// timing) and batches whose scratch exceeds FC_NOALLOC_SCRATCH_BYTES are
// rejected with a message.  Every call on a device shares that device's array:
// calls that overlap on one device race.
int World::relax_noalloc(int nstep_max, double Fconv, bool restore_start) {
    auto fail = [](const char* why) {
        std::fprintf(stderr, "[libMMFFmulti_lib] relax_noalloc: %s\n", why);
        return -1;
    };
    if (!dev_ready || nSys == 0 || !noalloc_ptr) return fail("not initialised");
    if (pc_mode == 2 && !learned_inkernel) return fail("the chunked learned form allocates; use relax()");
    if (pc_mode == 2 && !d_mlpW.p) return fail("learned gains need proposal weights");
    if (!ring_ok) return fail("asymmetric exclusions need the split kernels; use relax()");
    if (record_times) return fail("per-system timing allocates; switch it off (set_record_times(0))");
    if (autotune || h_order.empty()) plan_classes(false);
    if (n_unclassed > 0)
        return fail("periodic systems, substrate atoms and systems above 256 atoms need the split kernels; use relax()");
    // scratch layout: counters | order | activation slots | window cache (256 B aligned)
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t off_order = 256;
    const size_t off_hslot = off_order + up(h_order.size() * sizeof(int));
    std::vector<size_t> slot_off(classes.size(), 0);
    size_t nslot = 0;
    if (pc_mode == 2)
        for (size_t k = 0; k < classes.size(); k++) {
            slot_off[k] = nslot;
            nslot += (size_t)classes[k].ngroups * fc_hslot_elems(std::max(classes[k].nmax, 1));
        }
    const size_t off_cache = off_hslot + up(nslot * sizeof(unsigned short));
    const size_t need = off_cache + (size_t)ntotA * 16 * sizeof(float4);
    if (need > FC_NOALLOC_SCRATCH_BYTES) {
        std::fprintf(stderr,
                     "[libMMFFmulti_lib] relax_noalloc: the batch needs %zu bytes of scratch (%d atoms, %d systems), "
                     "more than the %u available; use relax()\n",
                     need, ntotA, nSys, (unsigned)FC_NOALLOC_SCRATCH_BYTES);
        return -1;
    }
    char* base = static_cast<char*>(noalloc_ptr);
    int* counters = reinterpret_cast<int*>(base);
    int* order = reinterpret_cast<int*>(base + off_order);
    unsigned short* hslot = reinterpret_cast<unsigned short*>(base + off_hslot);
    if (restore_start &&
        !hip_ok(hipMemcpyAsync(d_atoms.p, d_atoms0.p, sizeof(float4) * ntotV, hipMemcpyDeviceToDevice, stream),
                "restore"))
        return -1;
    if (!hip_ok(hipMemcpyAsync(order, h_order.data(), h_order.size() * sizeof(int), hipMemcpyHostToDevice, stream),
                "queue order") ||
        !hip_ok(fc_launch_zero_int(stream, counters, 8), "queue reset"))
        return -1;
    fcd::BatchDev B = batch_dev();
    B.systime = nullptr;
    B.gridcache = reinterpret_cast<float4*>(base + off_cache);
    B.gridcache_stride = ntotA;
    fc_launch_zero_f4(stream, d_stall.p, nSys);  // fallback flags of this relaxation only
    fcd::RunParams P = run_params(Fconv, nstep_max);
    hipEvent_t fork = noalloc_events[0];
    hipEventRecord(fork, stream);
    for (size_t k = 0; k < classes.size() && k < noalloc_events.size() - 1; k++) {
        const SizeClass& C = classes[k];
        hipStream_t st = class_streams[k % class_streams.size()];
        hipStreamWaitEvent(st, fork, 0);
        hipError_t err = fc_launch_relax_persistent(C.block, hb_batch ? 1 : 0, C.ngroups, C.lds, st, &B, &P, order,
                                                    qinfo.data(), counters, C.cls, C.nmax, C.mmax, C.wmax,
                                                    sched_mode == 0 ? 1 : 0, pc_mode == 2 ? hslot + slot_off[k] : nullptr,
                                                    C.h2lds);
        if (!hip_ok(err, "relax_noalloc launch")) return -1;
        hipEvent_t join = noalloc_events[1 + k];
        hipEventRecord(join, st);
        hipStreamWaitEvent(stream, join, 0);
    }
    d_nsteps.download(nsteps, stream);
    d_status.download(status, stream);
    if (!hip_ok(hipStreamSynchronize(stream), "relax_noalloc")) return -1;
    int mx = 0;
    for (int s = 0; s < nSys; s++) mx = std::max(mx, nsteps[s]);
    return mx;
}

// Screening scan (MolWorld_sp3_multi::scan_relaxed, MolWorld_sp3_multi.h:
// 1242-1296): configuration i is the template molecule moved by
// setFromRef(pos_i, rot_i) (atoms rot^T a + pos, pi-vectors rot^T p,
// MMFFsp3_loc.h:1543-1551) with atom 0 restrained to pos_i (constr.w = 1,
// constrK = 1, as the reference sets for its picked atom).  The reference
// relaxes nSystems configurations at a time and refills the slots that
// converged from the host; here all nconf configurations form one batch and
// the persistent kernel's work queue does the refill on the device: a
// workgroup takes the next configuration as soon as its system retires.
// Outputs per configuration: Es = force-field energy at the relaxed geometry
// (no restraint energy, as the reference's eval_no_omp), aforces = the
// force-field force on each atom (the restraint force removed), aposs = atom
// positions.  bRelax = false evaluates the placed configurations without
// relaxing them.  The batch that was loaded before is put back afterwards.
bool World::scan(int nconf, const double* poss, const double* rots, double* Es, double* aforces, double* aposs,
                 bool bRelax, int niter_max, double Fconv, double Flim) {
    if (!dev_ready || nSys == 0 || nconf <= 0) return false;
    // keep the current batch (device state included) to restore it afterwards
    if (!download_all(true, true)) return false;
    for (int s = 0; s < nSys; s++) unpack_system(s, true, true);
    std::vector<fc::MolSys> saved = mols;
    std::vector<float4> saved_constr = constr, saved_constrK = constrK, saved_bboxes = bboxes;
    const bool saved_grid = bGridFF;
    const fc::MolSys tmpl = has_templ ? templ : ffl;  // the template molecule of init
    const int N = tmpl.natoms, M = tmpl.nnode;
    std::vector<fc::MolSys> ms(nconf, tmpl);
    for (int i = 0; i < nconf; i++) {
        const double* p = poss ? poss + i * 3 : nullptr;
        const double* R = rots ? rots + i * 9 : nullptr;
        fc::MolSys& m = ms[i];
        for (int v = 0; v < N + M; v++) {
            const double* a = &tmpl.apos[v * 3];
            double o[3] = {a[0], a[1], a[2]};
            if (R)  // rot^T a  (Mat3::dot_to_T, rows of R in memory order)
                for (int q = 0; q < 3; q++) o[q] = R[q] * a[0] + R[3 + q] * a[1] + R[6 + q] * a[2];
            if (p && v < N)
                for (int q = 0; q < 3; q++) o[q] += p[q];
            for (int q = 0; q < 3; q++) m.apos[v * 3 + q] = o[q];
        }
        m.fapos.clear();
        m.vapos.clear();
    }
    set_systems(std::move(ms));  // the substrate grid (if any) stays loaded
    for (int i = 0; i < nconf; i++) {
        const int a0 = info[i].a0;
        const double* p = poss ? poss + i * 3 : nullptr;
        constr[a0] = make_float4(p ? (float)p[0] : 0.f, p ? (float)p[1] : 0.f, p ? (float)p[2] : 0.f, 1.f);
        constrK[a0] = make_float4(1.f, 1.f, 1.f, 1.f);
    }
    bool ok = upload_all(true, true, true);
    const float Flimit_keep = Flimit;
    if (Flim > 0) Flimit = (float)Flim;
    if (ok) ok = relax_persistent(bRelax ? niter_max : 0, Fconv, false) >= 0;
    Flimit = Flimit_keep;
    if (ok) ok = download_all(true, true);
    if (ok)
        for (int i = 0; i < nconf; i++) {
            const fcd::SysInfo& I = info[i];
            if (Es) Es[i] = Esys[i];
            for (int a = 0; a < N; a++) {
                const float4 x = atoms[I.v0 + a];
                float4 f = this->aforces[I.v0 + a];
                if (a == 0) {  // remove the restraint force of atom 0
                    const float4 c = constr[I.a0], k = constrK[I.a0];
                    f.x -= (c.x - x.x) * std::max(k.x, 0.f);
                    f.y -= (c.y - x.y) * std::max(k.y, 0.f);
                    f.z -= (c.z - x.z) * std::max(k.z, 0.f);
                }
                if (aforces) {
                    aforces[((size_t)i * N + a) * 3] = f.x;
                    aforces[((size_t)i * N + a) * 3 + 1] = f.y;
                    aforces[((size_t)i * N + a) * 3 + 2] = f.z;
                }
                if (aposs) {
                    aposs[((size_t)i * N + a) * 3] = x.x;
                    aposs[((size_t)i * N + a) * 3 + 1] = x.y;
                    aposs[((size_t)i * N + a) * 3 + 2] = x.z;
                }
            }
        }
    // the previous batch back
    set_systems(std::move(saved));
    constr = saved_constr;
    constrK = saved_constrK;
    bboxes = saved_bboxes;
    bGridFF = saved_grid;
    return upload_all(true, true, true) && ok;
}

// Systems that fit no workgroup class of the persistent kernel (N or M above
// 256, or N+M above 512) are relaxed after it by the split kernels, with plain
// FIRE on the exact force: the other systems keep their final state and are
// skipped.  Adds the device time to last_ms.
bool World::relax_unclassed(int nstep_max, double Fconv) {
    std::vector<int> only(nSys, 0);
    for (int s = 0; s < nSys; s++) only[s] = (sys_class[s] < 0) ? 1 : 0;
    d_only.upload(only, stream);
    fcd::BatchDev B = batch_dev();
    fcd::RunParams P = run_params(Fconv, nstep_max);
    P.pc_mode = 0;
    EventSet events;
    hipEvent_t e0 = events.make(), e1 = events.make();
    hipEventRecord(e0, stream);
    if (!hip_ok(fc_launch_reset_subset(stream, &B, &P, d_only.p), "reset subset")) return false;
    const int chunk = 50;
    for (int step = 0; step <= nstep_max;) {
        const int n = std::min(chunk, nstep_max + 1 - step);
        for (int k = 0; k < n; k++, step++)
            if (!hip_ok(fc_launch_split_step(stream, &B, &P, d_node_sys.p, ntotM, maxN, 1, step), "split step"))
                return false;
        d_status.download(status, stream);
        if (!hip_ok(hipStreamSynchronize(stream), "split status")) return false;
        bool any = false;
        for (int s = 0; s < nSys && !any; s++) any = (status[s] == fcd::ST_RUNNING);
        if (!any) break;
    }
    hipEventRecord(e1, stream);
    hipStreamSynchronize(stream);
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    last_ms += ms;
    return true;
}

int World::relax_split(int nstep_max, double Fconv, bool restore_start) {
    if (!dev_ready || nSys == 0) return -1;
    if (restore_start)
        hip_ok(hipMemcpyAsync(d_atoms.p, d_atoms0.p, sizeof(float4) * ntotV, hipMemcpyDeviceToDevice, stream),
               "restore");
    fcd::BatchDev B = batch_dev();
    fcd::RunParams P = run_params(Fconv, nstep_max);
    hipMemsetAsync(d_avel.p, 0, sizeof(float4) * ntotV, stream);
    fc_launch_zero_f4(stream, d_stall.p, nSys);  // no learned form here: no fallback flags
    fc_launch_reset(stream, &B, &P);
    EventSet events;
    hipEvent_t e0 = events.make(), e1 = events.make();
    hipEventRecord(e0, stream);
    const int chunk = 50;
    int step = 0;
    for (; step <= nstep_max;) {
        int n = std::min(chunk, nstep_max + 1 - step);
        for (int k = 0; k < n; k++, step++)
            if (!hip_ok(fc_launch_split_step(stream, &B, &P, d_node_sys.p, ntotM, maxN, 1, step), "split step"))
                return -1;
        d_status.download(status, stream);
        hip_ok(hipStreamSynchronize(stream), "split status");
        bool any = false;
        for (int s = 0; s < nSys && !any; s++) any = (status[s] == fcd::ST_RUNNING);
        if (!any) break;
    }
    hipEventRecord(e1, stream);
    hipStreamSynchronize(stream);
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    last_ms = ms;
    d_nsteps.download(nsteps, stream);
    d_status.download(status, stream);
    hip_ok(hipStreamSynchronize(stream), "split results");
    int mx = 0;
    for (int s = 0; s < nSys; s++) mx = std::max(mx, nsteps[s]);
    return mx;
}

// Global-optimisation exploring on the split kernels (the reference's
// run_ocl_opt with bGopt: updateMultiExploring every nPerVFs steps and the
// TDrive Langevin term in updateAtomsMMFFf4, MolWorld_sp3_multi.h:770-833,
// relax_multi.cl:1160-1178): every system alternates FIRE relaxation and
// Langevin exploring on the device (explore_move) for nstep_max steps and
// records the minima its relaxations reach.  Returns the total number of
// minima found, -1 on error.
int World::explore(int nstep_max, double Fconv, double T, double gamma, int n_explore, int n_relax,
                   unsigned seed, int kmax, bool keep_positions) {
    if (!dev_ready || nSys == 0) return -1;
    if (!std::isfinite(T) || !std::isfinite(gamma) || n_explore < 1 || kmax < 1 || nstep_max < 0) {
        std::fprintf(stderr, "[libMMFFmulti_lib] explore: needs finite T and gamma, nExplore >= 1, kmax >= 1\n");
        return -1;
    }
    // the reference drives only when gamma > 0 (TDrive.y > 0, relax_multi.cl:1155); T < 0 or gamma <= 0
    // leave the exploring phase undriven (constant-energy dynamics at dt_max)
    T = std::max(T, 0.0);
    gamma = std::max(gamma, 0.0);
    xkmax = kmax;
    d_xstate.upload(std::vector<int4>(nSys, make_int4(0, 0, 0, 0)), stream);
    d_minE.alloc((size_t)nSys * kmax);
    if (keep_positions) d_minpos.alloc((size_t)kmax * ntotV);
    else d_minpos.free_();
    fcd::BatchDev B = batch_dev();
    B.xstate = d_xstate.p;
    B.minE = d_minE.p;
    B.minpos = keep_positions ? d_minpos.p : nullptr;
    fcd::RunParams P = run_params(Fconv, nstep_max);
    P.pc_mode = 0;
    P.explore = 1;
    P.xT = (float)T;
    P.xgamma = (float)gamma;
    P.xn_explore = n_explore;
    P.xn_relax = std::max(n_relax, 0);
    P.xseed = seed;
    P.xkmax = kmax;
    hipMemsetAsync(d_avel.p, 0, sizeof(float4) * ntotV, stream);
    fc_launch_reset(stream, &B, &P);
    EventSet events;
    hipEvent_t e0 = events.make(), e1 = events.make();
    hipEventRecord(e0, stream);
    for (int step = 0; step <= nstep_max; step++)
        if (!hip_ok(fc_launch_split_step(stream, &B, &P, d_node_sys.p, ntotM, maxN, 1, step), "explore step"))
            return -1;
    hipEventRecord(e1, stream);
    // DevArr::download fills the host vector's current size: size them first
    std::vector<int4> xs((size_t)nSys);
    minE.assign((size_t)nSys * kmax, 0.0);
    minpos.assign(keep_positions ? (size_t)kmax * ntotV : 0, make_float4(0, 0, 0, 0));
    d_xstate.download(xs, stream);
    d_nsteps.download(nsteps, stream);
    d_status.download(status, stream);
    d_minE.download(minE, stream);
    if (keep_positions) d_minpos.download(minpos, stream);
    if (!hip_ok(hipStreamSynchronize(stream), "explore")) return -1;
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    last_ms = ms;
    nmin.assign(nSys, 0);
    int tot = 0;
    for (int s = 0; s < nSys; s++) {
        nmin[s] = xs[s].z;
        tot += xs[s].z;
    }
    minpos_kept = keep_positions;
    if (bSaveToDatabase && keep_positions) feed_database();
    return tot;
}

// MolWorld_sp3_multi::evalVFs with bSaveToDatabase (MolWorld_sp3_multi.h:863-879)
// offers every converged relaxation of an exploring run to the database as it
// happens; here the recorded minima are offered after the run, by record index
// and, within one index, in system order.
void World::feed_database() {
    std::vector<double> pos;
    for (int m = 0; m < xkmax; m++)
        for (int s = 0; s < nSys; s++) {
            if (m >= nmin[s]) continue;
            const fcd::SysInfo& I = info[s];
            const int N = I.natoms;
            pos.resize((size_t)N * 3);
            for (int i = 0; i < N; i++) {
                const float4 p = minpos[(size_t)m * ntotV + I.v0 + i];
                pos[i * 3] = p.x;
                pos[i * 3 + 1] = p.y;
                pos[i * 3 + 2] = p.z;
            }
            db.add_if_new(N, pos.data(), mols[s].elem.data(), minE[(size_t)s * xkmax + m], s);
        }
}

bool World::eval_forces() {
    if (!dev_ready || nSys == 0) return false;
    fcd::BatchDev B = batch_dev();
    fcd::RunParams P = run_params(0.0, 0);
    fc_launch_reset(stream, &B, &P);
    if (!hip_ok(fc_launch_split_step(stream, &B, &P, d_node_sys.p, ntotM, maxN, 0, 0), "eval_forces")) return false;
    return hip_ok(hipStreamSynchronize(stream), "eval_forces sync");
}

// Exact forces (split kernels) followed by the preconditioned force of the
// current gains into pcforces (host mirror, float4 per vec).
bool World::eval_precond() {
    if (!eval_forces()) return false;
    if (maxN > 256 || maxM > 256) {
        std::fprintf(stderr, "[libMMFFmulti_lib] eval_precond: systems above 256 atoms are not supported\n");
        return false;
    }
    fcd::BatchDev B = batch_dev();
    fcd::RunParams P = run_params(0.0, 0);
    d_pcforce.alloc(std::max(ntotV, 1));
    if (!hip_ok(fc_launch_precond(stream, &B, &P, d_pcforce.p, nullptr, nullptr), "precond")) return false;
    pcforces.resize(ntotV);
    d_pcforce.download(pcforces, stream);
    d_aforce.download(aforces, stream);
    return hip_ok(hipStreamSynchronize(stream), "eval_precond sync");
}

static unsigned short bf16_rne(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// Install proposal-network parameters (FP32, packing order of mlp_theta):
// dense weights rounded to bf16 for the matrix cores, the rest kept FP32.
bool World::set_mlp(const float* theta, int n) {
    const int NP = 64 * 16 + 64 + 64 * 128 + 64 + 64 * 128 + 64 + 2 * 64 + 2;
    if (n != NP) {
        std::fprintf(stderr, "[libMMFFmulti_lib] proposal weights: expected %d values, got %d\n", NP, n);
        return false;
    }
    for (int i = 0; i < n; i++)
        if (!std::isfinite(theta[i])) {
            std::fprintf(stderr, "[libMMFFmulti_lib] proposal weights: non-finite value at %d\n", i);
            return false;
        }
    if (!ensure_device()) return false;
    mlp_theta.assign(theta, theta + n);
    const float* W1 = theta;
    const float* b1 = W1 + 64 * 16;
    const float* W2 = b1 + 64;
    const float* b2 = W2 + 64 * 128;
    const float* W3 = b2 + 64;
    const float* b3 = W3 + 64 * 128;
    const float* W4 = b3 + 64;
    const float* b4 = W4 + 128;
    std::vector<unsigned short> wb;
    for (int i = 0; i < 64 * 16; i++) wb.push_back(bf16_rne(W1[i]));
    for (int i = 0; i < 64 * 128; i++) wb.push_back(bf16_rne(W2[i]));
    for (int i = 0; i < 64 * 128; i++) wb.push_back(bf16_rne(W3[i]));
    std::vector<float> fb;
    fb.insert(fb.end(), b1, b1 + 64);
    fb.insert(fb.end(), b2, b2 + 64);
    fb.insert(fb.end(), b3, b3 + 64);
    fb.insert(fb.end(), W4, W4 + 128);
    fb.insert(fb.end(), b4, b4 + 2);
    d_mlpW.upload(wb, stream);
    d_mlpF.upload(fb, stream);
    return hip_ok(hipStreamSynchronize(stream), "set_mlp");
}

bool World::sync() { return dev_ready ? hip_ok(hipStreamSynchronize(stream), "sync") : true; }

void World::register_buffers() {
    buffers.clear();
    fbuffers.clear();
    ibuffers.clear();
    // template system (MMFFmulti_lib.cpp:23-71): apos/pipos alias DOFs
    buffers["DOFs"] = DOFs.data();
    buffers["fDOFs"] = fDOFs.data();
    buffers["vDOFs"] = vDOFs.data();
    buffers["apos"] = DOFs.data();
    buffers["fapos"] = fDOFs.data();
    buffers["pipos"] = DOFs.data() + (size_t)ffl.natoms * 3;
    buffers["fpipos"] = fDOFs.data() + (size_t)ffl.natoms * 3;
    buffers["Es"] = Es;
    ibuffers["neighs"] = tmpl_neighs.data();
    ibuffers["ndims"] = ndims;
    ibuffers["selection"] = selection.data();
    ibuffers["gpu_neighs"] = (int*)neighs.data();
    ibuffers["gpu_neighCell"] = (int*)neighCell.data();
    ibuffers["gpu_bkNeighs"] = (int*)bkNeighs.data();
    fbuffers["gpu_atoms"] = (float*)atoms.data();
    fbuffers["gpu_aforces"] = (float*)aforces.data();
    fbuffers["gpu_avel"] = (float*)avel.data();
    fbuffers["gpu_constr"] = (float*)constr.data();
    fbuffers["gpu_constrK"] = (float*)constrK.data();
    fbuffers["gpu_bboxes"] = (float*)bboxes.data();
    fbuffers["gpu_REQs"] = (float*)REQs.data();
    fbuffers["gpu_MMpars"] = (float*)MMpars.data();
    fbuffers["gpu_BLs"] = (float*)BLs.data();
    fbuffers["gpu_BKs"] = (float*)BKs.data();
    fbuffers["gpu_Ksp"] = (float*)Ksp.data();
    fbuffers["gpu_Kpp"] = (float*)Kpp.data();
    fbuffers["gpu_lvecs"] = (float*)lvecs.data();
    fbuffers["gpu_ilvecs"] = (float*)ilvecs.data();
    fbuffers["gpu_pbcshifts"] = (float*)pbcshifts.data();
    // additive views of the uneven (CSR) batch
    ibuffers["sys_info"] = (int*)info.data();
    ibuffers["sys_nsteps"] = nsteps.data();
    ibuffers["sys_status"] = status.data();
    ibuffers["gpu_elem"] = elem.data();
    fbuffers["gpu_PLQs"] = (float*)PLQs.data();
    fbuffers["sys_Fmax"] = Fmax.data();
    pcforces.assign(std::max(ntotV, 1), make_float4(0.f, 0.f, 0.f, 0.f));
    fbuffers["gpu_pcforces"] = (float*)pcforces.data();
    ibuffers["gpu_comp"] = comp.data();
    buffers["sys_E"] = Esys.data();
}

}  // namespace fcw

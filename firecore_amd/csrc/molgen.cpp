// Synthetic molecule generator for the relaxation workloads.
//
// A molecule is a random branched tree of C/N/O centres (sp3 or sp2, at most
// 4 sigma neighbours) capped with H, laid out in the reference's
// MMFFsp3_loc conventions (nodes first, caps after, one pi-vector per node;
// parameters packed as MMFFBuilder::toMMFFsp3_loc does, MMFFBuilder.h:4153-4282).
// Parameter values follow the reference tables (AtomTypes.dat, BondTypes.dat)
// clamped to the workload's ranges (R0 1.2-1.9 A, E0 0.002-0.01 eV, bond
// lengths 1.09-1.50 A).
#include "fc_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace fc {

double Rng::normal() {
    if (has_spare) { has_spare = false; return spare; }
    double u, v, s;
    do {
        u = uniform(-1.0, 1.0);
        v = uniform(-1.0, 1.0);
        s = u * u + v * v;
    } while (s >= 1.0 || s == 0.0);
    double f = std::sqrt(-2.0 * std::log(s) / s);
    spare = v * f;
    has_spare = true;
    return u * f;
}

namespace {

struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalized(V3 a) { return a * (1.0 / norm(a)); }

// Node types.  Kss is stored x4 in apars.z (MMFFBuilder.h:4212); Ksp applies
// only to centres with a pi-orbital or lone pairs (MMFFBuilder.h:4243).
struct NType {
    int elem, deg, hyb;       // hyb: 3=sp3, 2=sp2
    double Ass, Asp, Kss, Ksp, Kpp, prob;
    bool sp2_parent_only;
};
const NType kTypes[] = {
    // elem  deg hyb  Ass    Asp    Kss   Ksp  Kpp  prob
    {EL_C, 4, 3, 109.5, -19.5, 6.32, 0.0, 0.0, 0.40, false},  // C sp3
    {EL_C, 3, 2, 120.0, 0.0, 6.34, 1.6, 1.3, 0.30, false},    // C sp2
    {EL_N, 3, 3, 109.5, -19.5, 5.60, 1.0, 0.5, 0.08, false},  // N sp3 (lone pair)
    {EL_N, 3, 2, 120.0, 0.0, 5.60, 1.6, 1.3, 0.07, false},    // N sp2
    {EL_O, 2, 3, 109.5, -19.5, 5.99, 1.0, 0.5, 0.08, false},  // O sp3 (lone pairs)
    {EL_O, 1, 2, 120.0, 0.0, 5.99, 1.6, 1.3, 0.07, true},     // O sp2 (carbonyl)
};
const int kNTypes = sizeof(kTypes) / sizeof(kTypes[0]);

// Per-element non-bonded parameters (R0 [A], E0 [eV]); REQ.y stores sqrt(E0)
// (MMFFparams.h:129).
const double kR0[4] = {1.9255, 1.8300, 1.7500, 1.4430};
const double kE0[4] = {0.00455323095, 0.00299212319, 0.00260184625, 0.0020};

void bond_params(int ea, int eb, bool sp2a, bool sp2b, double& l0, double& k) {
    if (ea == EL_H || eb == EL_H) {
        int e = (ea == EL_H) ? eb : ea;
        const double kH[3] = {16.457, 22.172, 24.436};
        l0 = 1.09;
        k = kH[e];
        return;
    }
    // symmetric heavy-atom table (single bonds)
    const double L[3][3] = {{1.500, 1.465, 1.423}, {1.465, 1.450, 1.420}, {1.423, 1.420, 1.450}};
    const double K[3][3] = {{10.082, 11.327, 12.723}, {11.327, 11.000, 11.500}, {12.723, 11.500, 10.000}};
    l0 = L[ea][eb];
    k = K[ea][eb];
    if (sp2a && sp2b) {  // conjugated / double
        l0 -= 0.12;
        k *= 1.8;
    }
}

// Canonical bond directions with slot 0 along -z; `lp` receives the
// pi / lone-pair direction.
void canonical_dirs(const NType& t, V3* d, V3& lp) {
    const double s8 = std::sqrt(8.0) / 3.0;
    V3 tet[4] = {{0, 0, -1},
                 {s8, 0, 1.0 / 3.0},
                 {s8 * std::cos(2 * M_PI / 3), s8 * std::sin(2 * M_PI / 3), 1.0 / 3.0},
                 {s8 * std::cos(4 * M_PI / 3), s8 * std::sin(4 * M_PI / 3), 1.0 / 3.0}};
    if (t.hyb == 3) {
        for (int k = 0; k < 4; k++) d[k] = tet[k];
        lp = tet[3];  // MMFFBuilder.h:4271 takes the 4th conformer direction
    } else {
        const double s3 = std::sqrt(3.0) / 2.0;
        d[0] = {0, 0, -1};
        d[1] = {s3, 0, 0.5};
        d[2] = {-s3, 0, 0.5};
        d[3] = {0, 1, 0};
        lp = {0, 1, 0};
    }
}

struct Frame {
    V3 e1, e2, e3;  // world images of canonical x, y, z
    V3 apply(V3 v) const { return e1 * v.x + e2 * v.y + e3 * v.z; }
};

// Frame whose canonical +z maps to `d` (so slot 0 points back along -d),
// rotated by psi about d.
Frame frame_along(V3 d, double psi) {
    V3 a = std::fabs(d.x) < 0.9 ? V3{1, 0, 0} : V3{0, 1, 0};
    V3 e1 = normalized(cross(d, a));
    V3 e2 = cross(d, e1);
    double c = std::cos(psi), s = std::sin(psi);
    return {e1 * c + e2 * s, e2 * c - e1 * s, d};
}

Frame random_frame(Rng& rng) {
    V3 d;
    do {
        d = {rng.normal(), rng.normal(), rng.normal()};
    } while (norm(d) < 1e-6);
    return frame_along(normalized(d), rng.uniform(0, 2 * M_PI));
}

struct GNode {
    int type, parent, pslot;
    V3 pos, lp;
    V3 dir[4];
    int nb[4];  // neighbour atom index per slot (filled later)
};

bool build_once(int natoms, Rng& rng, MolSys& out, double& min_nb_dist) {
    std::vector<GNode> nodes;
    std::vector<std::pair<int, int>> free_slots;  // (node, slot)
    auto allowed = [&](const NType& t, int maxinc, bool parent_sp2, bool allow_deg1) {
        if (t.deg - 1 > maxinc) return false;
        if (t.sp2_parent_only && !parent_sp2) return false;
        if (t.deg == 1 && !allow_deg1) return false;
        return true;
    };
    auto pick_type = [&](int maxinc, bool parent_sp2, bool allow_deg1) -> int {
        double tot = 0;
        for (int i = 0; i < kNTypes; i++)
            if (allowed(kTypes[i], maxinc, parent_sp2, allow_deg1)) tot += kTypes[i].prob;
        if (tot <= 0) return -1;
        double r = rng.uniform() * tot;
        for (int i = 0; i < kNTypes; i++) {
            if (!allowed(kTypes[i], maxinc, parent_sp2, allow_deg1)) continue;
            r -= kTypes[i].prob;
            if (r <= 0) return i;
        }
        return -1;
    };
    // occupied points: node positions and the probes of their open slots
    std::vector<V3> occ;
    std::vector<int> occ_owner;  // node id owning the point
    std::vector<int> occ_slot;   // -1 for the node itself, slot otherwise
    const double probe_len = 1.2;
    auto remove_probe = [&](int owner, int slot) {
        for (size_t o = 0; o < occ.size(); o++)
            if (occ_owner[o] == owner && occ_slot[o] == slot) {
                occ.erase(occ.begin() + o);
                occ_owner.erase(occ_owner.begin() + o);
                occ_slot.erase(occ_slot.begin() + o);
                return;
            }
    };
    auto clearance = [&](V3 x, int skip_owner, int skip_slot) {
        double d = 1e9;
        for (size_t o = 0; o < occ.size(); o++) {
            if (occ_owner[o] == skip_owner && (occ_slot[o] == -1 || occ_slot[o] == skip_slot)) continue;
            d = std::min(d, norm(x - occ[o]));
        }
        return d;
    };
    auto add_node_points = [&](int c) {
        const GNode& n = nodes[c];
        const NType& t = kTypes[n.type];
        occ.push_back(n.pos);
        occ_owner.push_back(c);
        occ_slot.push_back(-1);
        for (int k2 = (n.parent < 0 ? 0 : 1); k2 < t.deg; k2++) {
            occ.push_back(n.pos + n.dir[k2] * probe_len);
            occ_owner.push_back(c);
            occ_slot.push_back(k2);
        }
    };
    // root: a centre with >= 2 neighbours that fits
    int rt = -1;
    for (int tries = 0; tries < 64 && rt < 0; tries++) {
        int c = pick_type(natoms - 1, false, false);
        if (c >= 0 && kTypes[c].deg >= 2 && kTypes[c].deg + 1 <= natoms) rt = c;
    }
    if (rt < 0) return false;
    {
        GNode n{};
        n.type = rt;
        n.parent = -1;
        n.pslot = -1;
        V3 cd[4], clp;
        canonical_dirs(kTypes[rt], cd, clp);
        Frame F = random_frame(rng);
        n.pos = {0, 0, 0};
        for (int k2 = 0; k2 < 4; k2++) n.dir[k2] = F.apply(cd[k2]);
        n.lp = F.apply(clp);
        nodes.push_back(n);
        for (int k = 0; k < kTypes[rt].deg; k++) free_slots.push_back({0, k});
        add_node_points(0);
    }
    int total = 1 + kTypes[rt].deg;
    int guard = 0;
    while (total < natoms) {
        if (free_slots.empty() || ++guard > 100000) return false;
        const int need = natoms - total;
        // pick an open slot: random order, first with clearance >= 2.2 A at
        // ~1.45 A out, else the most open one
        int fi = -1;
        double best_cl = -1;
        const int ns = (int)free_slots.size();
        const int off = rng.randint(ns);
        for (int q = 0; q < ns; q++) {
            int k = (q + off) % ns;
            auto fs = free_slots[k];
            const GNode& p = nodes[fs.first];
            V3 x = p.pos + p.dir[fs.second] * 1.45;
            double cl = clearance(x, fs.first, fs.second);
            if (cl > best_cl) {
                best_cl = cl;
                fi = k;
            }
            if (cl >= 2.2) break;
        }
        auto fs = free_slots[fi];
        const GNode& pp = nodes[fs.first];
        const NType& pt = kTypes[pp.type];
        int t = pick_type(need, pt.hyb == 2, free_slots.size() > 2);
        if (t < 0) continue;
        free_slots.erase(free_slots.begin() + fi);
        GNode n{};
        n.type = t;
        n.parent = fs.first;
        n.pslot = fs.second;
        const NType& nt = kTypes[t];
        V3 cd[4], clp;
        canonical_dirs(nt, cd, clp);
        V3 d = pp.dir[n.pslot];
        double l0, kb;
        bond_params(pt.elem, nt.elem, pt.hyb == 2, nt.hyb == 2, l0, kb);
        n.pos = pp.pos + d * l0;
        std::vector<Frame> cands;
        if (nt.hyb == 2 && pt.hyb == 2) {
            V3 e2 = pp.lp;  // parent pi is perpendicular to d
            V3 e1 = cross(e2, d);
            cands.push_back({e1, e2, d});
            cands.push_back({e1 * -1.0, e2 * -1.0, d});
        } else {
            for (int q = 0; q < 12; q++) cands.push_back(frame_along(d, rng.uniform(0, 2 * M_PI)));
        }
        Frame best{};
        double best_score = -1;
        for (auto& F : cands) {
            double sc = 1e9;
            for (int k2 = 1; k2 < nt.deg; k2++)
                sc = std::min(sc, clearance(n.pos + F.apply(cd[k2]) * probe_len, n.parent, n.pslot));
            if (sc > best_score) {
                best_score = sc;
                best = F;
            }
        }
        for (int k2 = 0; k2 < 4; k2++) n.dir[k2] = best.apply(cd[k2]);
        n.lp = best.apply(clp);
        if (nt.deg == 1) n.lp = pp.lp;  // carbonyl pi follows its sp2 parent
        remove_probe(n.parent, n.pslot);
        const int id = (int)nodes.size();
        nodes.push_back(n);
        for (int k = 1; k < nt.deg; k++) free_slots.push_back({id, k});
        add_node_points(id);
        total += nt.deg - 1;
    }
    const int nnode = (int)nodes.size();
    const int ncap = natoms - nnode;
    if (ncap != (int)free_slots.size()) return false;

    // ---- atoms: nodes, then caps in (node, slot) order
    out.natoms = natoms;
    out.nnode = nnode;
    out.apos.assign((size_t)(natoms + nnode) * 3, 0.0);
    out.REQs.assign((size_t)natoms * 4, 0.0);
    out.neighs.assign((size_t)natoms * 4, -1);
    out.apars.assign((size_t)nnode * 4, 0.0);
    out.bLs.assign((size_t)nnode * 4, 0.0);
    out.bKs.assign((size_t)nnode * 4, 0.0);
    out.Ksp.assign((size_t)nnode * 4, 0.0);
    out.Kpp.assign((size_t)nnode * 4, 0.0);
    out.elem.assign((size_t)natoms, EL_H);
    std::vector<V3> pos(natoms);
    for (int c = 0; c < nnode; c++) {
        pos[c] = nodes[c].pos;
        out.elem[c] = kTypes[nodes[c].type].elem;
        for (int k = 0; k < 4; k++) nodes[c].nb[k] = -1;
    }
    for (int c = 1; c < nnode; c++) {
        nodes[c].nb[0] = nodes[c].parent;
        nodes[nodes[c].parent].nb[nodes[c].pslot] = c;
    }
    int icap = nnode;
    for (int c = 0; c < nnode; c++) {
        const NType& t = kTypes[nodes[c].type];
        for (int k = 0; k < t.deg; k++) {
            if (nodes[c].nb[k] >= 0) continue;
            double l0, kk;
            bond_params(t.elem, EL_H, t.hyb == 2, false, l0, kk);
            pos[icap] = nodes[c].pos + nodes[c].dir[k] * l0;
            nodes[c].nb[k] = icap;
            out.neighs[icap * 4 + 0] = c;
            icap++;
        }
    }
    for (int i = 0; i < natoms; i++) {
        out.apos[i * 3 + 0] = pos[i].x;
        out.apos[i * 3 + 1] = pos[i].y;
        out.apos[i * 3 + 2] = pos[i].z;
    }
    for (int c = 0; c < nnode; c++) {
        const GNode& n = nodes[c];
        const NType& t = kTypes[n.type];
        V3 lp = normalized(n.lp);
        out.apos[(natoms + c) * 3 + 0] = lp.x;
        out.apos[(natoms + c) * 3 + 1] = lp.y;
        out.apos[(natoms + c) * 3 + 2] = lp.z;
        double ang0 = t.Ass * M_PI / 180.0 * 0.5;
        out.apars[c * 4 + 0] = std::cos(ang0);
        out.apars[c * 4 + 1] = std::sin(ang0);
        out.apars[c * 4 + 2] = t.Kss * 4.0;
        out.apars[c * 4 + 3] = std::sin(t.Asp * M_PI / 180.0);
        for (int k = 0; k < 4; k++) {
            int j = n.nb[k];
            out.neighs[c * 4 + k] = j;
            if (j < 0) continue;
            bool jnode = j < nnode;
            int ej = out.elem[j];
            bool jsp2 = jnode && kTypes[nodes[j].type].hyb == 2;
            double l0, kb;
            bond_params(t.elem, ej, t.hyb == 2, jsp2, l0, kb);
            out.bLs[c * 4 + k] = l0;
            out.bKs[c * 4 + k] = kb;
            out.Ksp[c * 4 + k] = t.Ksp;  // zero for plain sp3 carbon
            double kppj = jnode ? kTypes[nodes[j].type].Kpp : 0.0;
            out.Kpp[c * 4 + k] = std::sqrt(t.Kpp * kppj);
        }
    }
    // non-bonded parameters and zero-sum charges
    double qsum = 0;
    std::vector<double> q(natoms);
    for (int i = 0; i < natoms; i++) {
        q[i] = 0.2 * rng.normal();
        qsum += q[i];
    }
    for (int i = 0; i < natoms; i++) {
        int e = out.elem[i];
        out.REQs[i * 4 + 0] = kR0[e];
        out.REQs[i * 4 + 1] = std::sqrt(kE0[e]);
        out.REQs[i * 4 + 2] = q[i] - qsum / natoms;
        out.REQs[i * 4 + 3] = 0.0;
    }
    // closest pair not related by 1-2 / 1-3
    min_nb_dist = 1e9;
    std::vector<std::vector<int>> adj(natoms);
    for (int i = 0; i < natoms; i++)
        for (int k = 0; k < 4; k++) {
            int j = out.neighs[i * 4 + k];
            if (j >= 0 && i < nnode) {
                adj[i].push_back(j);
                adj[j].push_back(i);
            }
        }
    for (int i = 0; i < natoms; i++) {
        std::vector<char> near(natoms, 0);
        near[i] = 1;
        for (int j : adj[i]) {
            near[j] = 1;
            for (int k : adj[j]) near[k] = 1;
        }
        for (int j = i + 1; j < natoms; j++)
            if (!near[j]) min_nb_dist = std::min(min_nb_dist, norm(pos[i] - pos[j]));
    }
    const double ax[3] = {1, 1, 1};
    flip_pis(out, ax);
    return true;
}

}  // namespace

bool gen_molecule(int natoms, uint64_t seed, MolSys& out) {
    if (natoms < 3) return false;
    MolSys best;
    double best_d = -1;
    for (int attempt = 0; attempt < 40; attempt++) {
        Rng rng(seed * 1000003ull + (uint64_t)attempt * 7919ull + (uint64_t)natoms);
        MolSys m;
        double d;
        if (!build_once(natoms, rng, m, d)) continue;
        if (d > best_d) {
            best_d = d;
            best = m;
        }
        if (d >= 1.9) break;
    }
    if (best_d < 0) return false;
    out = best;
    return true;
}

void perturb_positions(MolSys& m, uint64_t seed, double sigma) {
    Rng rng(seed ^ 0xA5A5A5A55A5A5A5Aull);
    for (int i = 0; i < m.natoms * 3; i++) m.apos[i] += sigma * rng.normal();
}

void flip_pis(MolSys& m, const double ax[3]) {
    for (int i = 0; i < m.nnode; i++) {
        double* p = &m.apos[(m.natoms + i) * 3];
        double c = p[0] * ax[0] + p[1] * ax[1] + p[2] * ax[2];
        if (c < 0) {
            p[0] = -p[0];
            p[1] = -p[1];
            p[2] = -p[2];
        }
    }
}

void place_molecule(MolSys& m, uint64_t seed, double ztarget, double cx, double cy) {
    Rng rng(seed ^ 0x5DEECE66Dull);
    // uniform random rotation from a unit quaternion
    double q[4];
    double n2;
    do {
        for (double& v : q) v = rng.normal();
        n2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    } while (n2 < 1e-12);
    double inv = 1.0 / std::sqrt(n2);
    for (double& v : q) v *= inv;
    double w = q[0], x = q[1], y = q[2], z = q[3];
    double R[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                   2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                   2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)};
    const int nv = m.nvecs();
    for (int i = 0; i < nv; i++) {
        double* p = &m.apos[i * 3];
        double a = R[0] * p[0] + R[1] * p[1] + R[2] * p[2];
        double b = R[3] * p[0] + R[4] * p[1] + R[5] * p[2];
        double c = R[6] * p[0] + R[7] * p[1] + R[8] * p[2];
        p[0] = a;
        p[1] = b;
        p[2] = c;
    }
    double zmin = 1e300, sx = 0, sy = 0;
    for (int i = 0; i < m.natoms; i++) {
        zmin = std::min(zmin, m.apos[i * 3 + 2]);
        sx += m.apos[i * 3 + 0];
        sy += m.apos[i * 3 + 1];
    }
    sx /= m.natoms;
    sy /= m.natoms;
    for (int i = 0; i < m.natoms; i++) {
        m.apos[i * 3 + 0] += cx - sx;
        m.apos[i * 3 + 1] += cy - sy;
        m.apos[i * 3 + 2] += ztarget - zmin;
    }
}

void merge_molecules(const MolSys& a, const MolSys& b, MolSys& out) {
    const int na = a.natoms, nb = b.natoms, ma = a.nnode, mb = b.nnode;
    const int ca = na - ma, cb = nb - mb;
    out.natoms = na + nb;
    out.nnode = ma + mb;
    const int N = out.natoms, M = out.nnode;
    // old -> new atom index maps
    auto mapA = [&](int i) { return i < 0 ? -1 : (i < ma ? i : ma + mb + (i - ma)); };
    auto mapB = [&](int i) { return i < 0 ? -1 : (i < mb ? ma + i : ma + mb + ca + (i - mb)); };
    out.apos.assign((size_t)(N + M) * 3, 0);
    out.REQs.assign((size_t)N * 4, 0);
    out.neighs.assign((size_t)N * 4, -1);
    out.elem.assign(N, EL_H);
    for (int i = 0; i < na; i++) {
        int j = mapA(i);
        std::memcpy(&out.apos[j * 3], &a.apos[i * 3], 3 * sizeof(double));
        std::memcpy(&out.REQs[j * 4], &a.REQs[i * 4], 4 * sizeof(double));
        for (int k = 0; k < 4; k++) out.neighs[j * 4 + k] = mapA(a.neighs[i * 4 + k]);
        out.elem[j] = a.elem[i];
    }
    for (int i = 0; i < nb; i++) {
        int j = mapB(i);
        std::memcpy(&out.apos[j * 3], &b.apos[i * 3], 3 * sizeof(double));
        std::memcpy(&out.REQs[j * 4], &b.REQs[i * 4], 4 * sizeof(double));
        for (int k = 0; k < 4; k++) out.neighs[j * 4 + k] = mapB(b.neighs[i * 4 + k]);
        out.elem[j] = b.elem[i];
    }
    (void)cb;
    for (int i = 0; i < ma; i++) std::memcpy(&out.apos[(N + i) * 3], &a.apos[(na + i) * 3], 3 * sizeof(double));
    for (int i = 0; i < mb; i++) std::memcpy(&out.apos[(N + ma + i) * 3], &b.apos[(nb + i) * 3], 3 * sizeof(double));
    auto cat = [](const std::vector<double>& x, const std::vector<double>& y) {
        std::vector<double> r(x);
        r.insert(r.end(), y.begin(), y.end());
        return r;
    };
    out.apars = cat(a.apars, b.apars);
    out.bLs = cat(a.bLs, b.bLs);
    out.bKs = cat(a.bKs, b.bKs);
    out.Ksp = cat(a.Ksp, b.Ksp);
    out.Kpp = cat(a.Kpp, b.Kpp);
}

// System i of a synthetic workload (see init_workload in capi.cpp).
//  kind 1: one molecule, natoms ~ U[nmin,nmax], seed seed0+i
//  kind 2: two molecules ~ U[nmin,nmax] each, random orientations, centroids
//          4-8 A apart (pushed apart until the closest pair is >= 2.5 A)
// Every system gets N(0,sigma) noise; with a substrate it is rotated at random
// and placed with its lowest atom 3.0 +- 0.3 A above the grid origin plane and
// its centroid at a random lateral point of the cell.
bool make_workload_system(int kind, int i, long long seed0, int nmin, int nmax, double sigma, bool surf,
                          const double* pos0, double ax, double by, MolSys& m) {
    const uint64_t sd = (uint64_t)(seed0 + i);
    Rng rng(sd * 2654435761ull + 17);
    if (kind == 1) {
        int n = nmin + rng.randint(nmax - nmin + 1);
        if (!gen_molecule(n, sd, m)) return false;
    } else {
        MolSys a, b;
        int na = nmin + rng.randint(nmax - nmin + 1);
        int nb = nmin + rng.randint(nmax - nmin + 1);
        if (!gen_molecule(na, sd * 2 + 0x1000000ull, a) || !gen_molecule(nb, sd * 2 + 0x1000001ull, b)) return false;
        place_molecule(a, sd * 3 + 1, 0.0, 0.0, 0.0);
        place_molecule(b, sd * 3 + 2, 0.0, 0.0, 0.0);
        double dir[3] = {rng.normal(), rng.normal(), rng.normal()};
        double nd = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        for (double& x : dir) x /= nd;
        double sep = rng.uniform(4.0, 8.0);
        auto centroid = [](const MolSys& x, double* c) {
            c[0] = c[1] = c[2] = 0;
            for (int k = 0; k < x.natoms; k++)
                for (int q = 0; q < 3; q++) c[q] += x.apos[k * 3 + q] / x.natoms;
        };
        double ca[3], cb[3];
        centroid(a, ca);
        centroid(b, cb);
        for (int it = 0; it < 200; it++) {
            double shift[3];
            for (int q = 0; q < 3; q++) shift[q] = ca[q] + dir[q] * sep - cb[q];
            double dmin = 1e9;
            for (int x = 0; x < a.natoms; x++)
                for (int y = 0; y < b.natoms; y++) {
                    double d2 = 0;
                    for (int q = 0; q < 3; q++) {
                        double d = a.apos[x * 3 + q] - (b.apos[y * 3 + q] + shift[q]);
                        d2 += d * d;
                    }
                    dmin = std::min(dmin, d2);
                }
            if (std::sqrt(dmin) >= 2.5 || it == 199) {
                for (int y = 0; y < b.natoms; y++)
                    for (int q = 0; q < 3; q++) b.apos[y * 3 + q] += shift[q];
                break;
            }
            sep += 0.25;
        }
        merge_molecules(a, b, m);
    }
    perturb_positions(m, sd + 7777777ull, sigma);
    if (surf)
        place_molecule(m, sd + 31337ull, pos0[2] + 3.0 + rng.uniform(-0.3, 0.3), pos0[0] + rng.uniform(0, ax),
                       pos0[1] + rng.uniform(0, by));
    return true;
}

}  // namespace fc

// Typed molecule from an .xyz file: the reference's MMFF build chain of
// MolWorld_sp3::init for xyz_name with bMMFF (MolWorld_sp3.h:309-357), i.e.
//   initParams            type tables (MMFFparams::init, MMFFparams.h:531-542)
//   buildMolecule_xyz     loadGeom (MolWorld_sp3.h:1143-1180): atoms typed by
//                         name (MMFFBuilder::loadXYZ_Atoms, MMFFBuilder.h:
//                         3725-3764), bonds by covalent radii (autoBonds
//                         :2087-2115, autoBondsPBC :2120-2166 when the file
//                         has an lvs line), then lone pairs and pi-bond counts
//                         (autoConfEPi :1632-1691) and bond parameters from
//                         the pi counts (assignBondParams :867-892)
//   makeMMFFs             nodes first (sortConfAtomsFirst :2448-2458), bond
//                         atoms ordered (checkBondsOrdered :2333-2343), atom
//                         types refined (assignTypes :1444-1452 with
//                         assignAllSp3Types :1232-1244 and
//                         assignSpecialTypes :1297-1428), MMFFsp3_loc
//                         parameters and pi-vectors (toMMFFsp3_loc
//                         :4153-4282, makeConfGeom :942-1008), caps'
//                         neighbours (makeBackNeighs), flipPis((1,1,1))
//   makeFFs               non-bonded parameters of the final types
//                         (initNBmol -> MMFFparams::assignREs :783-794,
//                         charges from the file, H-bond term 0)
// without electron pairs (bEpairs = false, the ctypes module's default).
// Everything here is host code of the library; the result is one MolSys in
// the layout the batch packer takes.
#include "fc_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <unordered_map>
#include <unordered_set>

namespace fc {

namespace {

bool skip_line(const std::string& line) {
    if (!line.empty() && line[0] == '#') return true;
    return line.find_first_not_of(" \t\r\n") == std::string::npos;
}

// whitespace-separated fields, as the tables' sscanf formats read them
std::vector<std::string> fields(const std::string& line) {
    std::vector<std::string> f;
    std::istringstream ss(line);
    std::string t;
    while (ss >> t) f.push_back(t);
    return f;
}

// sscanf %i / %lf of one field; false when the field does not start a number
bool as_int(const std::string& s, int& v) {
    char* e = nullptr;
    long x = std::strtol(s.c_str(), &e, 0);
    if (e == s.c_str()) return false;
    v = (int)x;
    return true;
}
bool as_double(const std::string& s, double& v) {
    char* e = nullptr;
    double x = std::strtod(s.c_str(), &e);
    if (e == s.c_str()) return false;
    v = x;
    return true;
}
// number of leading fields that parse (sscanf stops at the first failure)
template <class F>
int count_parsed(const std::vector<std::string>& f, int n, F ok) {
    int k = 0;
    while (k < n && k < (int)f.size() && ok(k, f[k])) k++;
    return k;
}

}  // namespace

int TypeTables::atom_type(const std::string& name) const {
    auto it = adict.find(name);
    return it == adict.end() ? -1 : it->second;
}

// MMFFparams::loadElementTypes / loadAtomTypes / loadBondTypes and
// assignAllSubTypes (MMFFparams.h:196-345, 346-384, 469-523).
bool TypeTables::load(const std::string& fElem, const std::string& fAtom, const std::string& fBond,
                      std::string& err) {
    etypes.clear();
    edict.clear();
    atypes.clear();
    adict.clear();
    bonds.clear();
    {
        std::ifstream in(fElem);
        if (!in) return err = "cannot read " + fElem, false;
        std::string line;
        while (std::getline(in, line)) {
            if (skip_line(line)) continue;
            const auto f = fields(line);
            // name iZ neval valence piMax color(hex) Rcov RvdW EvdW Quff Uuff Vuff [QEq x4]
            ElementTypeB e;
            const int n = count_parsed(f, 12, [&](int k, const std::string& s) {
                int iv;
                double dv;
                switch (k) {
                    case 0: e.name = s; return true;
                    case 1: return as_int(s, e.iZ);
                    case 2: return as_int(s, e.neval);
                    case 3: return as_int(s, e.valence);
                    case 4: return as_int(s, e.piMax);
                    case 5: return std::sscanf(s.c_str(), "%x", (unsigned*)&iv) == 1;
                    case 6: return as_double(s, e.Rcov);
                    case 7: return as_double(s, e.RvdW);
                    case 8: return as_double(s, e.EvdW);
                    default: return as_double(s, dv);
                }
            });
            if (n < 12) return err = "incomplete element type line in " + fElem + ": " + line, false;
            if (!edict.emplace(e.name, (int)etypes.size()).second)
                return err = "duplicate element type " + e.name, false;
            etypes.push_back(e);
        }
    }
    {
        std::ifstream in(fAtom);
        if (!in) return err = "cannot read " + fAtom, false;
        std::string line;
        while (std::getline(in, line)) {
            if (skip_line(line)) continue;
            const auto f = fields(line);
            // name parent element epair valence nepair npi sym Ruff RvdW EvdW Qbase Hb [Ass Asp Kss Ksp Kep Kpp]
            AtomTypeB t;
            std::string parent, element, epair;
            int iv[4] = {0, 0, 0, 0};
            double dv[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            const int n = count_parsed(f, 19, [&](int k, const std::string& s) {
                switch (k) {
                    case 0: t.name = s; return true;
                    case 1: parent = s; return true;
                    case 2: element = s; return true;
                    case 3: epair = s; return true;
                    default:
                        if (k < 8) return as_int(s, iv[k - 4]);
                        return as_double(s, dv[k - 8]);
                }
            });
            if (n < 13) return err = "incomplete atom type line in " + fAtom + ": " + line, false;
            // the reference reads these with %i into one-byte fields: they keep the low byte
            t.valence = (uint8_t)iv[0];
            t.nepair = (uint8_t)iv[1];
            t.npi = (uint8_t)iv[2];
            t.sym = (uint8_t)iv[3];
            t.Ruff = dv[0];
            t.RvdW = dv[1];
            t.EvdW = dv[2];
            t.Qbase = dv[3];
            t.Hb = dv[4];
            t.bMMFF = n >= 19;
            if (t.bMMFF) {
                t.Ass = dv[5];
                t.Asp = dv[6];
                t.Kss = dv[7];
                t.Ksp = dv[8];
                t.Kep = dv[9];
                t.Kpp = dv[10];
            }
            if (!atypes.empty()) {  // the first type (the wildcard) resolves nothing
                const auto ie = edict.find(element);
                const int ipar = atom_type(parent), iept = atom_type(epair);
                if (ie == edict.end() || ipar < 0 || iept < 0)
                    return err = "atom type " + t.name + " refers to an unknown element, parent or lone-pair type", false;
                t.element = ie->second;
                t.parent = ipar;
                t.ePairType = iept;
                t.iZ = etypes[ie->second].iZ;
            }
            if (!adict.emplace(t.name, (int)atypes.size()).second) return err = "duplicate atom type " + t.name, false;
            atypes.push_back(t);
        }
        // sub-types X_3, X_2, X_1 of the first type of each atomic number
        std::unordered_set<int> done;
        for (AtomTypeB& t : atypes) {
            if (!done.insert(t.iZ).second) continue;
            const char* ssub[3] = {"3", "2", "1"};
            for (int i = 0; i < 3; i++) {
                const int it = atom_type(t.name + "_" + ssub[i]);
                if (it >= 0) t.subTypes[i] = it;
            }
        }
    }
    {
        std::ifstream in(fBond);
        if (!in) return err = "cannot read " + fBond, false;
        std::string line;
        while (std::getline(in, line)) {
            if (skip_line(line)) continue;
            const auto f = fields(line);
            int order = 0;
            double l0 = 0, k = 0;
            if (f.size() < 5 || !as_int(f[2], order) || !as_double(f[3], l0) || !as_double(f[4], k))
                return err = "incomplete bond type line in " + fBond + ": " + line, false;
            const int a = atom_type(f[0]), b = atom_type(f[1]);
            if (a < 0 || b < 0) return err = "bond type of unknown atom types: " + line, false;
            bonds.emplace(bond_key(a, b, order), BondTypeB{l0, k});  // the first of duplicates stays
        }
    }
    return true;
}

uint64_t TypeTables::bond_key(int a, int b, int order) {
    // BondType::getId (MMFFparams.h:32): 16-bit type indices sorted, 8-bit order
    uint16_t x = (uint16_t)a, y = (uint16_t)b;
    if (x > y) std::swap(x, y);
    return ((uint64_t)x) | ((uint64_t)y << 16) | ((uint64_t)(uint8_t)order << 32);
}

// MMFFparams::getBondParams -> getBondType (MMFFparams.h:596-618, 743-747):
// the exact pair, then the pair of the element types (the type named as the
// element), else the defaults 2.0 A and 10.0 eV/A^2.
void TypeTables::bond_params(int ti, int tj, int order, double& l0, double& k) const {
    auto it = bonds.find(bond_key(ti, tj, order));
    if (it == bonds.end()) {
        const int i0 = atom_type(etypes[atypes[ti].element].name);
        const int j0 = atom_type(etypes[atypes[tj].element].name);
        if (i0 >= 0 && j0 >= 0) it = bonds.find(bond_key(i0, j0, order));
    }
    if (it == bonds.end()) {
        l0 = 2.0;
        k = 10.0;
        return;
    }
    l0 = it->second.l0;
    k = it->second.k;
}

namespace {

struct V {
    double x = 0, y = 0, z = 0;
};
inline V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V operator*(V a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V cross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
// Vec3::normalize (Vec3.h:149-154): multiply by the reciprocal of the norm
inline double normalize(V& v) {
    const double n = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    const double inv = 1.0 / n;
    v.x *= inv;
    v.y *= inv;
    v.z *= inv;
    return n;
}
// Vec3::getSomeOrtho (Vec3.h:220-245): two vectors orthogonal to v (not normalised)
inline void some_ortho(V v, V& v1, V& v2) {
    const double xx = v.x * v.x, yy = v.y * v.y;
    if (xx < yy) {
        v1 = {-yy - v.z * v.z, v.x * v.y, v.x * v.z};
    } else {
        v1 = {v.y * v.x, -v.z * v.z - xx, v.y * v.z};
    }
    v2 = {v.y * v1.z - v.z * v1.y, v.z * v1.x - v.x * v1.z, v.x * v1.y - v.y * v1.x};
}

// MMFFBuilder::makeConfGeom (MMFFBuilder.h:942-1008): the remaining directions
// hs[nb..3] around an atom from its nb bond directions and npi pi-bonds; hs[3]
// is the pi-vector.
void conf_geom(int nb, int npi, V* hs) {
    V ma, mb, mc;
    if (nb == 3) {
        mb = cross(hs[1] - hs[0], hs[2] - hs[0]);
        mb = mb * (-1 / std::sqrt(dot(mb, mb)));
        if (npi == 0 && 0 < dot(mb, hs[0] + hs[1] + hs[2])) mb = mb * -1.;
        hs[3] = mb;
    } else if (nb == 2) {
        // Mat3::fromCrossSafe (Mat3.h:385-395)
        mb = cross(hs[0], hs[1]);
        ma = hs[1] - hs[0];
        normalize(ma);
        const double r2b = dot(mb, mb);
        if (r2b < 1e-15) {
            some_ortho(ma, mb, mc);
        } else {
            mb = mb * (1 / std::sqrt(r2b));
            mc = cross(mb, ma);
        }
        if (npi == 0) {
            const double cb = 0.81649658092, cc = 0.57735026919;
            hs[2] = mc * cc + mb * cb;
            hs[3] = mc * cc - mb * cb;
        } else {
            hs[2] = mc;
            hs[3] = mb;
        }
    } else if (nb == 1) {
        mc = hs[0];
        normalize(mc);
        some_ortho(mc, mb, ma);
        if (npi == 0) {
            const double ca = 0.81649658092, cb = 0.47140452079, cc = -0.33333333333;
            hs[1] = mc * cc + mb * (cb * 2);
            hs[2] = mc * cc - mb * cb + ma * ca;
            hs[3] = mc * cc - mb * cb - ma * ca;
        } else if (npi == 1) {
            const double ca = 0.86602540378, cc = -0.5;
            hs[1] = mc * cc + ma * ca;
            hs[2] = mc * cc - ma * ca;
            hs[3] = mb;
        } else {
            hs[1] = mc * -1;
            hs[2] = mb;
            hs[3] = ma;
        }
    } else if (nb == 0) {
        mc = hs[0];
        normalize(mc);
        some_ortho(mc, mb, ma);
        if (npi == 0) {
            const double ca = 0.81649658092, cb = 0.47140452079, cc = -0.33333333333;
            hs[0] = mc * cc + mb * (cb * 2);
            hs[1] = mc * cc - mb * cb + ma * ca;
            hs[2] = mc * cc - mb * cb - ma * ca;
            hs[3] = mc;
        }
    }
}

struct BAtom {
    int type;
    V pos;
    double REQ[4];
    int iconf;
};
struct BConf {
    int iatom;
    int nbond = 0, npi = 0, ne = 0, nH = 0, n = 0;
    int neighs[4] = {-1, -1, -1, -1};  // bond indices
};
struct BBond {
    int a, b;
    double l0 = 1.0, k = 0.0;
};

}  // namespace

bool build_from_xyz(const std::string& fname, const TypeTables& T, bool& periodic, double* lvec9, MolSys& out,
                    std::string& err) {
    // ---- MMFFparams::loadXYZ (MMFFparams.h:844-896)
    std::ifstream in(fname);
    if (!in) return err = "cannot read " + fname, false;
    std::string line;
    if (!std::getline(in, line)) return err = "empty file " + fname, false;
    const int na = std::atoi(line.c_str());
    if (na <= 0) return err = "no atom count in " + fname, false;
    if (!std::getline(in, line)) return err = "no comment line in " + fname, false;
    periodic = false;
    {  // cellFromString: the line must start with exactly "lvs" and hold 9 numbers
        double L[9];
        if (line.size() >= 3 && line.compare(0, 3, "lvs") == 0 &&
            std::sscanf(line.c_str() + 3, "%lf %lf %lf %lf %lf %lf %lf %lf %lf", L, L + 1, L + 2, L + 3, L + 4, L + 5,
                        L + 6, L + 7, L + 8) == 9) {
            periodic = true;
            std::memcpy(lvec9, L, sizeof(L));
        }
    }
    const int iE = T.atom_type("E"), iH = T.atom_type("H");
    std::unordered_set<int> capping;  // types of atomic number 1 and 200 (addCappingTypesByIz)
    for (int i = 0; i < (int)T.atypes.size(); i++)
        if (T.atypes[i].iZ == 1 || T.atypes[i].iZ == 200) capping.insert(i);
    std::vector<BAtom> atoms;
    std::vector<BConf> confs;
    for (int i = 0; i < na; i++) {
        if (!std::getline(in, line)) return err = "unexpected end of " + fname, false;
        char name[64];
        double x, y, z, Q = 0, H = 0;
        const int n = std::sscanf(line.c_str(), "%63s %lf %lf %lf %lf %lf", name, &x, &y, &z, &Q, &H);
        if (n < 4) return err = "incomplete atom line in " + fname + ": " + line, false;
        if (n < 5) Q = 0;
        if (n < 6) H = 0;
        const int ityp = T.atom_type(name);
        if (ityp < 0) return err = std::string("unknown atom type ") + name + " in " + fname, false;
        // MMFFBuilder::loadXYZ_Atoms + insertAtom(ityp, p, REQ, npi=0, ne)
        BAtom A;
        A.type = ityp;
        A.pos = {x, y, z};
        A.REQ[0] = T.atypes[ityp].RvdW;
        A.REQ[1] = std::sqrt(T.atypes[ityp].EvdW);
        A.REQ[2] = Q;
        A.REQ[3] = H;
        A.iconf = -1;
        if (!capping.count(ityp)) {
            BConf c;
            c.iatom = (int)atoms.size();
            c.ne = T.atypes[ityp].nepair;
            c.n = c.ne;
            A.iconf = (int)confs.size();
            confs.push_back(c);
        }
        atoms.push_back(A);
    }
    (void)iH;
    // ---- bonds by covalent radii: autoBonds (R = -1.20) or autoBondsPBC (R = -1.35,
    // images -1..1 of the bonded partner, ix outer, iz inner)
    std::vector<BBond> bonds;
    auto add_bond = [&](int i, int j) -> bool {
        const int ib = (int)bonds.size();
        bonds.push_back(BBond{i, j});
        for (int a : {i, j}) {  // tryAddBondToAtomConf -> AtomConf::addBond
            const int ic = atoms[a].iconf;
            if (ic < 0) continue;
            BConf& c = confs[ic];
            if (c.n >= 4) return false;
            c.neighs[c.nbond] = ib;
            c.nbond++;
            c.n++;
        }
        return true;
    };
    {
        const double Rfac = periodic ? 1.35 : 1.20;
        const int np = periodic ? 1 : 0;
        const V la{lvec9[0], lvec9[1], lvec9[2]}, lb{lvec9[3], lvec9[4], lvec9[5]}, lc{lvec9[6], lvec9[7], lvec9[8]};
        for (int i = 0; i < na; i++) {
            const BAtom& A = atoms[i];
            const bool capi = capping.count(A.type) > 0;
            const double Ri = T.atypes[A.type].Ruff;
            for (int ix = -np; ix <= np; ix++)
                for (int iy = -np; iy <= np; iy++)
                    for (int iz = -np; iz <= np; iz++) {
                        // Mat3::lincomb(ix, iy, iz) = a*ix + b*iy + c*iz
                        const V p = periodic ? A.pos - (la * (double)ix + lb * (double)iy + lc * (double)iz) : A.pos;
                        for (int j = i + 1; j < na; j++) {  // touchingAtoms
                            const V dp = atoms[j].pos - p;
                            const double R = (T.atypes[atoms[j].type].Ruff + Ri) * Rfac;
                            if (!(dot(dp, dp) < R * R)) continue;
                            if (capi && capping.count(atoms[j].type)) continue;
                            if (!add_bond(i, j))
                                return err = "atom " + std::to_string(i) + " or " + std::to_string(j) +
                                             " has more than 4 bonds and lone pairs",
                                       false;
                        }
                    }
        }
    }
    // ---- checkNumberOfBonds (MMFFBuilder.h:2209-2237, called with bCheckInit at
    // MolWorld_sp3.h:1173, where the reference exits): every atom bonded, a node
    // within its type's valence, a capping atom with one bond
    {
        std::vector<int> nb(na, 0);
        for (const BBond& b : bonds) {
            nb[b.a]++;
            nb[b.b]++;
        }
        for (int ia = 0; ia < na; ia++) {
            const std::string at = "atom " + std::to_string(ia) + " (" + T.atypes[atoms[ia].type].name + ")";
            if (nb[ia] <= 0) return err = at + " has no bonds", false;
            if (atoms[ia].iconf >= 0) {
                if (nb[ia] > T.atypes[atoms[ia].type].valence)
                    return err = at + " has " + std::to_string(nb[ia]) + " bonds, more than its valence", false;
            } else if (nb[ia] > 1) {
                return err = "capping " + at + " has " + std::to_string(nb[ia]) + " bonds", false;
            }
        }
    }
    // ---- autoConfEPi (MMFFBuilder.h:1632-1691): lone pairs of the type, pi count
    // = 4 - bonds - lone pairs - caps, at most the element's piMax
    for (int ia = 0; ia < na; ia++) {
        const int ic = atoms[ia].iconf;
        if (ic < 0) continue;
        BConf& c = confs[ic];
        const int ne = T.atypes[atoms[ia].type].nepair;
        if (c.nbond + c.nH + ne > 4)
            return err = "atom " + std::to_string(ia) + ": bonds + lone pairs exceed 4", false;
        c.ne = ne;
        c.npi = 4 - c.nbond - c.ne - c.nH;
        c.n = 4;
        c.npi = std::min(c.npi, T.etypes[T.atypes[atoms[ia].type].element].piMax);
    }
    // ---- assignAllBondParams: order 1 + min(npi_i, npi_j) (both with confs),
    // parameters of the current (element-named) types
    for (BBond& b : bonds) {
        int order = 1;
        const int ci = atoms[b.a].iconf, cj = atoms[b.b].iconf;
        if (ci >= 0 && cj >= 0) order += std::min(confs[ci].npi, confs[cj].npi);
        T.bond_params(atoms[b.a].type, atoms[b.b].type, order, b.l0, b.k);
    }
    // ---- sortConfAtomsFirst: atoms with a conf first, both groups in file order
    {
        std::vector<int> permut(na);
        int j = 0;
        for (int i = 0; i < na; i++)
            if (atoms[i].iconf >= 0) permut[i] = j++;
        for (int i = 0; i < na; i++)
            if (atoms[i].iconf < 0) permut[i] = j++;
        for (BBond& b : bonds) {
            b.a = permut[b.a];
            b.b = permut[b.b];
        }
        std::vector<BAtom> old = atoms;
        for (int i = 0; i < na; i++) {
            atoms[permut[i]] = old[i];
            if (old[i].iconf >= 0) confs[old[i].iconf].iatom = permut[i];
        }
        for (BBond& b : bonds)  // checkBondsOrdered(bOrder = true)
            if (!(b.a < b.b)) std::swap(b.a, b.b);
    }
    auto partner = [&](int ib, int ia) { return bonds[ib].a == ia ? bonds[ib].b : (bonds[ib].b == ia ? bonds[ib].a : -1); };
    // ---- assignTypes: sp3/sp2/sp1 sub-types of C, N, O, then the special types
    {
        std::unordered_set<int> sp3types;
        for (const char* s : {"C", "N", "O"}) {
            const int t = T.atom_type(s);
            if (t >= 0) sp3types.insert(t);
        }
        for (BAtom& A : atoms) {  // assignAllSp3Types
            if (T.atypes[A.type].parent != 0) continue;
            if (A.iconf < 0) continue;
            if (!sp3types.count(A.type)) continue;
            A.type = T.atypes[A.type].subTypes[confs[A.iconf].npi];
        }
        std::vector<int> ng((size_t)na * 4, -1);  // makeNeighs
        for (int ia = 0; ia < na; ia++) {
            const int ic = atoms[ia].iconf;
            if (ic < 0) continue;
            const BConf& c = confs[ic];
            for (int k = 0; k < c.nbond; k++) {
                const int ib = c.neighs[k];
                if (ib < 0) continue;
                const int ja = partner(ib, ia);
                ng[ia * 4 + k] = ja;
                if (atoms[ja].iconf == -1) ng[ja * 4] = ia;
            }
        }
        auto ty = [&](const char* s) {
            const int t = T.atom_type(s);
            return t;
        };
        const int iC_2 = ty("C_2"), iC_R = ty("C_R"), iC_COO = ty("C_COO"), iC_ald = ty("C_ald"), iC_3 = ty("C_3"),
                  iC_ene = ty("C_ene"), iC_yne = ty("C_yne"), iC_CH3 = ty("C_CH3");
        const int iO_3 = ty("O_3"), iO_2 = ty("O_2"), iO_OH = ty("O_OH"), iO_ald = ty("O_ald"), iO_sCOO = ty("O_sCOO"),
                  iO_pCOO = ty("O_pCOO");
        const int iN_3 = ty("N_3"), iN_2 = ty("N_2"), iN_R = ty("N_R");
        const int iH_OH = ty("H_OH"), iH_NH2 = ty("H_NH2"), iH_CH3 = ty("H_CH3"), iH_ene = ty("H_ene"),
                  iH_yne = ty("H_yne"), iH_ald = ty("H_ald");
        for (int t : {iC_2, iC_R, iC_COO, iC_ald, iC_3, iC_ene, iC_yne, iC_CH3, iO_3, iO_2, iO_OH, iO_ald, iO_sCOO,
                      iO_pCOO, iN_3, iN_2, iN_R, iH_OH, iH_NH2, iH_CH3, iH_ene, iH_yne, iH_ald})
            if (t < 0) return err = "the atom type table lacks a type of the special-type rules", false;
        auto count_type = [&](int ia, int t) {  // hasNeighborOfType
            int n = 0;
            for (int k = 0; k < 4; k++) {
                const int ja = ng[ia * 4 + k];
                if (ja >= 0 && atoms[ja].type == t) n++;
            }
            return n;
        };
        int itr = 0;
        for (; itr < 10; itr++) {  // assignSpecialTypesLoop
            int nnew = 0;
            for (int ia = 0; ia < na; ia++) {
                BAtom& A = atoms[ia];
                const int iZ = T.atypes[A.type].iZ;
                int npineigh = 0;  // countNeighPi
                if (A.iconf >= 0)
                    for (int k = 0; k < 4; k++) {
                        const int ja = ng[ia * 4 + k];
                        if (ja < 0 || atoms[ja].iconf < 0) continue;
                        if (confs[atoms[ja].iconf].npi) npineigh++;
                    }
                int itnew = A.type;
                if (iZ == 1) {
                    const int ja = ng[ia * 4];
                    const int ingt = (ja < 0 || ja > na) ? -1 : atoms[ja].type;
                    if (ingt >= 0) {
                        if (ingt == iO_OH) itnew = iH_OH;
                        else if (ingt == iN_3) itnew = iH_NH2;
                        else if (ingt == iN_R) itnew = iH_NH2;
                        else if (ingt == iC_3) itnew = iH_CH3;
                        else if (ingt == iC_CH3) itnew = iH_CH3;
                        else if (ingt == iC_2) itnew = iH_ene;
                        else if (ingt == iC_R) itnew = iH_ene;
                        else if (ingt == iC_ene) itnew = iH_ene;
                        else if (ingt == iC_yne) itnew = iH_yne;
                        else if (ingt == iC_ald) itnew = iH_ald;
                    }
                } else if (iZ == 6) {
                    if (A.type == iC_2) {
                        const bool hasO3 = A.iconf >= 0 && count_type(ia, iO_3) > 0;
                        const bool hasO2 = A.iconf >= 0 && count_type(ia, iO_2) > 0;
                        if (hasO3 && hasO2) itnew = iC_COO;
                        else if (hasO2) itnew = iC_ald;
                        else if (npineigh >= 2) itnew = iC_R;
                    }
                } else if (iZ == 7) {
                    if (A.type == iN_3 && npineigh >= 1) itnew = iN_R;
                } else if (iZ == 8) {
                    const bool hasCOO = A.iconf >= 0 && count_type(ia, iC_COO) > 0;
                    const bool hasAld = A.iconf >= 0 && count_type(ia, iC_ald) > 0;
                    if (hasCOO) {
                        if (A.type == iO_3) itnew = iO_sCOO;
                        else if (A.type == iO_2) itnew = iO_pCOO;
                    } else if (hasAld) {
                        itnew = iO_ald;
                    }
                }
                if (itnew >= 0 && itnew != A.type) {
                    A.type = itnew;
                    nnew++;
                }
            }
            if (nnew == 0) break;
        }
        if (itr == 10) return err = "special atom types did not settle in 10 passes", false;
        (void)iN_2;
    }
    // ---- toMMFFsp3_loc (bEPairs = false): nodes = atoms with a conf
    int nnode = 0;
    for (const BAtom& A : atoms) nnode += A.iconf >= 0 ? 1 : 0;
    MolSys& m = out;
    m = MolSys{};
    m.natoms = na;
    m.nnode = nnode;
    m.apos.assign((size_t)(na + nnode) * 3, 0.0);
    m.REQs.assign((size_t)na * 4, 0.0);
    m.neighs.assign((size_t)na * 4, -1);
    m.apars.assign((size_t)nnode * 4, 0.0);
    m.bLs = m.bKs = m.Ksp = m.Kpp = m.apars;
    m.elem.assign(na, EL_C);
    const double deg2rad = 0.01745329251994329547;  // cpp/common/constants.h:6
    for (int ia = 0; ia < na; ia++) {
        const BAtom& A = atoms[ia];
        m.apos[ia * 3] = A.pos.x;
        m.apos[ia * 3 + 1] = A.pos.y;
        m.apos[ia * 3 + 2] = A.pos.z;
        const AtomTypeB& at = T.atypes[A.type];
        // non-bonded parameters of the final type (assignREs, bSqrtE, bQ0 = false,
        // bHB0 = true), charge from the file
        m.REQs[ia * 4] = at.RvdW;
        m.REQs[ia * 4 + 1] = std::sqrt(at.EvdW);
        m.REQs[ia * 4 + 2] = A.REQ[2];
        m.REQs[ia * 4 + 3] = 0.0;
        const int iZ = at.iZ;
        m.elem[ia] = iZ == 1 ? EL_H : (iZ == 7 ? EL_N : (iZ == 8 ? EL_O : EL_C));
        if (A.iconf < 0) continue;
        const BConf& c = confs[A.iconf];
        if (c.npi > 2) return err = "atom " + std::to_string(ia) + " has more than 2 pi-bonds", false;
        const double ang0 = at.Ass * deg2rad * 0.5;
        m.apars[ia * 4] = std::cos(ang0);
        m.apars[ia * 4 + 1] = std::sin(ang0);
        m.apars[ia * 4 + 2] = at.Kss * 4.0;
        m.apars[ia * 4 + 3] = std::sin(at.Asp * deg2rad);
        V hs[4];
        for (int k = 0; k < c.nbond; k++) {
            const BBond& B = bonds[c.neighs[k]];
            const int ja = partner(c.neighs[k], ia);
            const AtomTypeB& jt = T.atypes[atoms[ja].type];
            hs[k] = atoms[ja].pos - A.pos;
            normalize(hs[k]);
            m.neighs[ia * 4 + k] = ja;
            m.bLs[ia * 4 + k] = B.l0;
            m.bKs[ia * 4 + k] = B.k;
            m.Ksp[ia * 4 + k] = (c.npi > 0 || c.ne > 0) ? at.Ksp : 0.0;
            m.Kpp[ia * 4 + k] = std::sqrt(at.Kpp * jt.Kpp);
        }
        conf_geom(c.nbond, c.npi, hs);
        // pi-vector = hs[3], then flipPis((1,1,1)); the init's check evaluation
        // (MolWorld_sp3.h:1558-1562, eval() -> normalizePis, MMFFsp3_loc.h:1037)
        // leaves it normalised
        V p = hs[3];
        if (dot(p, V{1, 1, 1}) < 0) p = p * -1;
        normalize(p);
        m.apos[(na + ia) * 3] = p.x;
        m.apos[(na + ia) * 3 + 1] = p.y;
        m.apos[(na + ia) * 3 + 2] = p.z;
    }
    // caps' neighbours from back-neighbours (makeBackNeighs with bCapNeighs)
    std::vector<int> bk((size_t)na * 4);
    if (!make_back_neighs(na, nnode, m.neighs.data(), bk.data()))
        return err = "an atom has more than 4 back-neighbours", false;
    for (int ia = nnode; ia < na; ia++) {
        if (bk[ia * 4] < 0) return err = "cap atom " + std::to_string(ia) + " has no bond", false;
        m.neighs[ia * 4] = bk[ia * 4] / 4;
        m.neighs[ia * 4 + 1] = m.neighs[ia * 4 + 2] = m.neighs[ia * 4 + 3] = -1;
    }
    (void)iE;
    return true;
}

}  // namespace fc

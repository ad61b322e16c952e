// Host-side data structures of the multi-system MMFF relaxation library.
//
// One relaxation "system" is one molecule (nodes first, caps after, then one
// pi-vector per node), stored in the layout of MMFFsp3_loc
// (cpp/common/molecular/MMFFsp3_loc.h:40-130 and MMFFBuilder.h:4153-4282):
//   vecs   [0,nnode)            node atoms (<=4 bonded neighbours + pi-vector)
//          [nnode,natoms)       cap atoms (one neighbour)
//          [natoms,natoms+nnode) pi-vectors of the nodes
//   REQs   (R0, sqrt(E0), Q, H) per atom
//   neighs 4 local atom indices per atom, -1 padded (caps: {node,-1,-1,-1})
//   apars  (cos(Ass/2), sin(Ass/2), 4*Kss, sin(Asp)) per node
//   bLs, bKs, Ksp, Kpp: 4 per node (one per neighbour slot)
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

namespace fc {

enum Elem { EL_C = 0, EL_N = 1, EL_O = 2, EL_H = 3 };

struct MolSys {
    int natoms = 0;
    int nnode = 0;
    std::vector<double> apos;   // (natoms+nnode)*3
    std::vector<double> REQs;   // natoms*4
    std::vector<int> neighs;    // natoms*4
    std::vector<double> apars;  // nnode*4
    std::vector<double> bLs;    // nnode*4
    std::vector<double> bKs;    // nnode*4
    std::vector<double> Ksp;    // nnode*4
    std::vector<double> Kpp;    // nnode*4
    std::vector<int> elem;      // natoms (EL_*)
    std::vector<double> fapos;  // (natoms+nnode)*3, filled by unpack with forces
    std::vector<double> vapos;  // (natoms+nnode)*3, filled by unpack with velocities
    // periodic molecule (MMFFsp3_loc with bPBC): lattice vectors as rows a, b, c
    // and the image range -nPBC..nPBC per axis; vacuum when pbc is false
    bool pbc = false;
    double lvec[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int nPBC[3] = {0, 0, 0};
    int nvecs() const { return natoms + nnode; }
};

// ---------------- deterministic RNG (splitmix64 + Box-Muller) --------------
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    int randint(int n) { return (int)(uniform() * n); }
    double normal();
    bool has_spare = false;
    double spare = 0;
};

// ---------------- generator (molgen.cpp) ----------------
// Random branched tree of C/N/O sp3/sp2 centres with H caps, exactly `natoms`
// atoms, ideal geometry (no noise).  Returns false when natoms is not
// reachable (natoms < 3).
bool gen_molecule(int natoms, uint64_t seed, MolSys& out);
// Adds N(0, sigma) noise to every atom position (pi-vectors untouched).
void perturb_positions(MolSys& m, uint64_t seed, double sigma);
// Random rigid rotation, then translation so that the lowest atom sits at
// z = ztarget and the centroid at (cx, cy).
void place_molecule(MolSys& m, uint64_t seed, double ztarget, double cx, double cy);
// Flip pi-vectors into the half-space of `ax` (MMFFsp3_loc.h:1167-1172).
void flip_pis(MolSys& m, const double ax[3]);
// Pair of molecules in one system (config 4): concatenated topology.
void merge_molecules(const MolSys& a, const MolSys& b, MolSys& out);
// System i of a synthetic workload (kind 1: single molecules, kind 2: pairs).
bool make_workload_system(int kind, int i, long long seed0, int nmin, int nmax, double sigma, bool surf,
                          const double* pos0, double ax, double by, MolSys& m);

// ---------------- topology (topology.cpp) ----------------
// makeBackNeighs (MMFFsp3_loc.h:1335-1358): bk[j] lists ia*4+k for every node
// ia with neighs[ia][k]==j, filled first-empty.  Returns false on >4.
bool make_back_neighs(int natoms, int nnode, const int* neighs, int* bkneighs);
// Initial unit pi-vectors of the nodes from the geometry (MMFFsp3_loc::initPi,
// MMFFsp3_loc.h:882-945, non-periodic): pipos (nnode x 3).  flip: then turn
// every pi-vector into the half-space of (1,1,1) (flipPis, :1167-1172).
void init_pi(int natoms, int nnode, const double* apos, const int* neighs, const double* Ksp, const double* Kpp,
             bool flip, double* pipos);
// Non-bonded pair weights in the bNonBondNeighs strategy:
//   w(i,j) = [j not bonded to i] - (#nodes k whose angle loop subtracts {i,j})
// Pairs with w==1 need nothing; every other pair (including i==j) is marked in
// the exclusion bit-matrix (row i, bit j).  Pairs with w not in {0,1} are also
// returned in `corr` as (i,j,w) so that w*pair(i,j) can be added back.
struct PairCorr { int i, j; float w; };
void make_exclusions(int natoms, int nnode, const int* neighs, bool subtract13,
                     std::vector<uint32_t>& mask, int& words_per_row,
                     std::vector<PairCorr>& corr);
// Rotated rows (bit k of row i = pair (i, (i+k) mod N)) for the half-ring
// schedule of the persistent kernel; false if the exclusions are asymmetric.
int connected_components(int natoms, const int* neighs, int maxc, int* comp);
// Periodic molecules: image shifts (MolWorld_sp3::evalPBCshifts order) and the
// image of each bonded neighbour (MMFFsp3_loc::makeNeighCells(npbc, shifts)).
int make_pbc_shifts(const double* lvec, const int* nPBC, std::vector<double>& shifts);
void make_neigh_cells(int natoms, const double* apos, const int* neighs, int npbc, const double* shifts,
                      int* neighCell);
bool rotate_exclusions(int natoms, const std::vector<uint32_t>& mask, int W, std::vector<uint32_t>& rot);

// ---------------- typed molecules from .xyz (builder.cpp) ----------------
// Parameter tables of the reference (ElementTypes.dat, AtomTypes.dat,
// BondTypes.dat; MMFFparams.h:59-130, 196-384).
struct ElementTypeB {
    std::string name;
    int iZ = 0, neval = 0, valence = 0, piMax = 0;
    double Rcov = 0, RvdW = 0, EvdW = 0;
};
struct AtomTypeB {
    std::string name;
    int iZ = 0;
    uint8_t valence = 0, nepair = 0, npi = 0, sym = 0;
    double Ruff = 0, RvdW = 0, EvdW = 0, Qbase = 0, Hb = 0;
    int parent = 0, element = 0, ePairType = 0;
    bool bMMFF = false;
    double Ass = 0, Asp = 0, Kss = 0, Ksp = 0, Kep = 0, Kpp = 0;
    int subTypes[3] = {0, 0, 0};  // the type's _3, _2, _1 variants (0 when absent)
};
struct BondTypeB {
    double l0, k;
};
struct TypeTables {
    std::vector<ElementTypeB> etypes;
    std::unordered_map<std::string, int> edict;
    std::vector<AtomTypeB> atypes;
    std::unordered_map<std::string, int> adict;
    std::map<uint64_t, BondTypeB> bonds;
    bool load(const std::string& fElem, const std::string& fAtom, const std::string& fBond, std::string& err);
    int atom_type(const std::string& name) const;  // -1 if unknown
    static uint64_t bond_key(int a, int b, int order);
    void bond_params(int ti, int tj, int order, double& l0, double& k) const;
};
// The reference's MMFF build of an .xyz molecule (MolWorld_sp3::init with
// bMMFF, bEpairs = false): typed atoms, bonds, MMFFsp3_loc parameters and
// pi-vectors, nodes first.  periodic/lvec9 report an lvs line.  False with a
// message in err on a file or typing error.
bool build_from_xyz(const std::string& fname, const TypeTables& T, bool& periodic, double* lvec9, MolSys& out,
                    std::string& err);

// ---------------- minima database of the exploring mode (database.cpp) ----------------
// descriptor of a structure (principal axes + atom count): returns its length
int db_descriptor(int n, const double* pos, double* d4);
// sqrt(sum over y's atoms of the squared distance to the nearest same-type atom
// of x after the optimal superposition of the centred x on the centred y)
double db_distance(int n, const double* x, const int* tx, const double* y, const int* ty);
struct MinimaDB {
    static constexpr double DESC_TOL = 5.0;  // descriptor distance of candidates
    static constexpr double DIST_TOL = 0.1;  // db_distance / natoms of duplicates
    struct Member {
        int natoms = 0, ndesc = 0, isys = -1, occurrence = 0;
        double desc[4] = {0, 0, 0, 0};
        double E = 0;
        std::vector<double> pos;
        std::vector<int> types;
    };
    std::vector<Member> members;
    long total = 0;  // structures offered
    // index of the member the structure duplicates (its count raised), or -1
    // when it was added as a new member
    int add_if_new(int n, const double* pos, const int* types, double E, int isys);
    void clear() {
        members.clear();
        total = 0;
    }
};

// ---------------- GridFF (gridff_host.cpp) ----------------
struct GridFFHost {
    int nx = 0, ny = 0, nz = 0;
    double pos0[3] = {0, 0, 0};
    double step[3] = {0.1, 0.1, 0.1};   // orthogonal grid spacing
    double alphaMorse = 1.5;
    std::vector<double> coefs;          // (nx,ny,nz,3) C-order, z fastest
};
bool npy_load_f64(const std::string& fname, std::vector<int>& shape, std::vector<double>& data);
bool npy_save_f64(const std::string& fname, const std::vector<int>& shape, const double* data);
// Surface atoms: positions (n*3) and REQs (n*4) with a 2D lattice (a, b).
struct Surface {
    std::vector<double> apos, REQs;
    double lvec[9] = {0};
    // Coulomb channel of the node sums: 0 = bare k*q/r (the reference's
    // evalGridFFPoint_Coul, r^2 bounded below by GRIDFF_R2MIN); > 0 = damped
    // k*q/sqrt(r^2 + Rdamp^2), the form of the reference's GPU Coulomb
    // (getLJQH with R2damp, relax_multi.cl:151-164)
    double Rdamp = 0.0;
    int natoms() const { return (int)(apos.size() / 3); }
};
void make_nacl_slab(double a, int nlayers, double q, Surface& s);
bool load_surface_xyz(const std::string& fname, Surface& s);
bool save_surface_xyz(const std::string& fname, const Surface& s);
// GridFF generation (gridff_host.cpp): node potentials (Pauli, London,
// Coulomb) as the reference's GridFF::evalBsplineRef, then exact cubic
// B-spline interpolation (periodic in x,y, open in z).
constexpr double GRIDFF_R2MIN = 1e-4;  // (0.01 A)^2: lower bound of r^2 in the bare Coulomb node sums
// Coulomb damping of the synthetic NaCl(001) substrate of the benchmark
// ("synth:NaCl"): its molecules carry random charges (hydrogens up to
// ~0.5 e), which a bare ionic potential pulls onto the top-layer ions and,
// below the grid's first nodes, through the slab
constexpr double NACL_SYNTH_RDAMP = 1.0;
void gridff_images(const Surface& s, int& nmx, int& nmy, int& ncx, int& ncy);
void gridff_point(const Surface& s, const double* p, double alpha, int nmx, int nmy, int ncx, int ncy, double* out);
void gridff_geometry(const Surface& s, int nx, int ny, int nz, double dz, double z0, double alpha, GridFFHost& g);
void gridff_nodes_host(const Surface& s, const GridFFHost& g, int nthreads, std::vector<double>& V);
void gridff_fit(GridFFHost& g, std::vector<double>& V, int nthreads);

}  // namespace fc

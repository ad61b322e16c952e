// Topology bookkeeping: back-neighbours and non-bonded exclusions.
#include "fc_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace fc {

// Back-neighbour lists, following MMFFsp3_loc::makeBackNeighs
// (cpp/common/molecular/MMFFsp3_loc.h:1335-1358): for every node ia and slot
// j with neighs[ia][j]=ja>=0, append ia*4+j to the first empty entry of
// bkneighs[ja].  Cap atoms are not scanned (their own neighbour lists point
// back to their node).
bool make_back_neighs(int natoms, int nnode, const int* neighs, int* bk) {
    for (int i = 0; i < natoms * 4; i++) bk[i] = -1;
    for (int ia = 0; ia < nnode; ia++) {
        for (int j = 0; j < 4; j++) {
            int ja = neighs[ia * 4 + j];
            if (ja < 0) continue;
            if (ja >= natoms) return false;
            int* row = bk + ja * 4;
            int k = 0;
            while (k < 4 && row[k] >= 0) k++;
            if (k == 4) return false;
            row[k] = ia * 4 + j;
        }
    }
    return true;
}

// Pair weights of the non-bonded sum in the "bNonBondNeighs" strategy of the
// reference CPU path:
//  * NBFF::evalLJQs_ng4_atom_omp (NBFF.h:666-691) skips j==i and j in
//    neighs[i] (1-2 exclusion by the neighbour list);
//  * MMFFsp3_loc::eval_atom (MMFFsp3_loc.h:466-552) subtracts the full pair
//    interaction once for every pair of neighbours of a node (1-3) when
//    doAngles && bSubtractAngleNonBond.
// The net weight of the directed pair i->j is therefore
//    w = [j not in neighs[i]] - n13(i,j)
// which is 0 for plain 1-2 and 1-3 pairs of a tree and 1 for all other pairs.
void make_exclusions(int natoms, int nnode, const int* neighs, bool subtract13,
                     std::vector<uint32_t>& mask, int& W, std::vector<PairCorr>& corr) {
    W = (natoms + 31) / 32;
    mask.assign((size_t)natoms * W, 0u);
    corr.clear();
    std::vector<int> w((size_t)natoms * natoms, 1);
    for (int i = 0; i < natoms; i++) {
        w[(size_t)i * natoms + i] = 0;
        for (int k = 0; k < 4; k++) {
            int j = neighs[i * 4 + k];
            if (j >= 0 && j < natoms) w[(size_t)i * natoms + j] = 0;
        }
    }
    if (subtract13) {
        for (int k = 0; k < nnode; k++) {
            const int* ng = neighs + k * 4;
            for (int a = 0; a < 3; a++) {
                int i = ng[a];
                if (i < 0) break;
                for (int b = a + 1; b < 4; b++) {
                    int j = ng[b];
                    if (j < 0) break;
                    w[(size_t)i * natoms + j] -= 1;
                    w[(size_t)j * natoms + i] -= 1;
                }
            }
        }
    }
    for (int i = 0; i < natoms; i++) {
        for (int j = 0; j < natoms; j++) {
            int wij = w[(size_t)i * natoms + j];
            if (wij == 1) continue;
            mask[(size_t)i * W + (j >> 5)] |= (1u << (j & 31));
            if (i != j && wij != 0) corr.push_back({i, j, (float)wij});
        }
    }
}

// Rotated exclusion rows for the half-ring pair schedule: bit k of row i is
// the exclusion bit of the pair (i, (i+k) mod N).  Returns false when the
// exclusions are not symmetric (then every pair needs both directions).
bool rotate_exclusions(int natoms, const std::vector<uint32_t>& mask, int W, std::vector<uint32_t>& rot) {
    rot.assign((size_t)natoms * W, 0u);
    bool sym = true;
    for (int i = 0; i < natoms; i++) {
        for (int k = 0; k < natoms; k++) {
            const int j = (i + k) % natoms;
            const bool b = (mask[(size_t)i * W + (j >> 5)] >> (j & 31)) & 1u;
            const bool bt = (mask[(size_t)j * W + (i >> 5)] >> (i & 31)) & 1u;
            if (b != bt) sym = false;
            if (b || bt) rot[(size_t)i * W + (k >> 5)] |= 1u << (k & 31);
        }
    }
    return sym;
}

// Connected components of the bonded graph given by the neighbour lists
// (bonds listed on either end).  Component indices are assigned in order of
// the lowest atom index; indices >= maxc are folded into maxc-1 (the
// molecule-level projection then treats those molecules as one rigid group).
// Returns the number of distinct (folded) indices.
int connected_components(int natoms, const int* neighs, int maxc, int* comp) {
    std::vector<int> parent(natoms);
    for (int i = 0; i < natoms; i++) parent[i] = i;
    auto find = [&](int x) {
        while (parent[x] != x) x = parent[x] = parent[parent[x]];
        return x;
    };
    for (int i = 0; i < natoms; i++)
        for (int k = 0; k < 4; k++) {
            const int j = neighs[i * 4 + k];
            if (j < 0 || j >= natoms) continue;
            int a = find(i), b = find(j);
            if (a != b) parent[std::max(a, b)] = std::min(a, b);
        }
    std::vector<int> id(natoms, -1);
    int nc = 0;
    for (int i = 0; i < natoms; i++) {
        const int r = find(i);
        if (id[r] < 0) id[r] = nc++;
        comp[i] = std::min(id[r], maxc - 1);
    }
    return std::min(nc, maxc);
}

}  // namespace fc

namespace fc {

// Pi-vectors of the nodes, MMFFsp3_loc::initPi (MMFFsp3_loc.h:882-945) for a
// molecule without periodic images.  First pass: a node whose pi-sigma
// stiffness Ksp is above Kmin for at least two neighbours takes the unit
// normal of the plane through the first two such neighbours and the third
// (or the node itself when there are only two), kept when its squared length
// before normalisation exceeds r2min.  Second pass, over the nodes still
// without a pi-vector in node order: the pi-vector of the neighbour with the
// largest Kpp (strictly largest first found), if that Kpp exceeds Kmin, else
// +z.  Then optionally flipPis(1,1,1) (:1167-1172).
void init_pi(int natoms, int nnode, const double* apos, const int* neighs, const double* Ksp, const double* Kpp,
             bool flip, double* pipos) {
    (void)natoms;
    const double Kmin = 1e-4, r2min = 1e-4;
    for (int ia = 0; ia < nnode; ia++) {
        const int* ng = neighs + ia * 4;
        const double* ks = Ksp + ia * 4;
        double pts[3][3];
        int nfound = 0;
        for (int j = 0; j < 4 && nfound < 3; j++) {
            if (ks[j] > Kmin) {
                for (int q = 0; q < 3; q++) pts[nfound][q] = apos[ng[j] * 3 + q];
                nfound++;
            }
        }
        double* out = pipos + ia * 3;
        out[0] = out[1] = out[2] = 0.0;
        if (nfound < 2) continue;
        const double* p = (nfound == 2) ? apos + ia * 3 : pts[2];
        const double u[3] = {pts[0][0] - p[0], pts[0][1] - p[1], pts[0][2] - p[2]};
        const double v[3] = {pts[1][0] - p[0], pts[1][1] - p[1], pts[1][2] - p[2]};
        const double c[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        const double r2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
        if (r2 > r2min) {
            const double s = 1.0 / std::sqrt(r2);
            for (int q = 0; q < 3; q++) out[q] = c[q] * s;
        }
    }
    for (int ia = 0; ia < nnode; ia++) {
        double* out = pipos + ia * 3;
        if (out[0] * out[0] + out[1] * out[1] + out[2] * out[2] > 0.1) continue;
        const int* ng = neighs + ia * 4;
        const double* kp = Kpp + ia * 4;
        int imax = -1;
        double kmax = 0.0;
        for (int i = 0; i < 4; i++)
            if (kp[i] > kmax) {
                imax = i;
                kmax = kp[i];
            }
        if (kmax > Kmin && ng[imax] >= 0 && ng[imax] < nnode) {
            const double* src = pipos + ng[imax] * 3;
            for (int q = 0; q < 3; q++) out[q] = src[q];
        } else {
            out[0] = 0.0;
            out[1] = 0.0;
            out[2] = 1.0;
        }
    }
    if (flip)
        for (int ia = 0; ia < nnode; ia++) {
            double* out = pipos + ia * 3;
            if (out[0] + out[1] + out[2] < 0.0)
                for (int q = 0; q < 3; q++) out[q] = -out[q];
        }
}

// Periodic images (MolWorld_sp3::evalPBCshifts, MolWorld_sp3.h:756-765): shift
// ipbc = ix*a + iy*b + iz*c for iz outer, iy, ix inner over -nPBC..nPBC.
int make_pbc_shifts(const double* lvec, const int* nPBC, std::vector<double>& shifts) {
    shifts.clear();
    for (int iz = -nPBC[2]; iz <= nPBC[2]; iz++)
        for (int iy = -nPBC[1]; iy <= nPBC[1]; iy++)
            for (int ix = -nPBC[0]; ix <= nPBC[0]; ix++)
                for (int q = 0; q < 3; q++) shifts.push_back(lvec[q] * ix + lvec[3 + q] * iy + lvec[6 + q] * iz);
    return (int)(shifts.size() / 3);
}

// Image of each bonded neighbour (MMFFsp3_loc::makeNeighCells(npbc, shifts),
// MMFFsp3_loc.h:1394-1434): the first image with the smallest |pj + shift -
// pi|^2, compared in single precision as the reference does; -1 for empty slots.
void make_neigh_cells(int natoms, const double* apos, const int* neighs, int npbc, const double* shifts,
                      int* neighCell) {
    for (int ia = 0; ia < natoms; ia++)
        for (int j = 0; j < 4; j++) {
            const int ja = neighs[ia * 4 + j];
            neighCell[ia * 4 + j] = -1;
            if (ja < 0) continue;
            const double d[3] = {apos[ja * 3] - apos[ia * 3], apos[ja * 3 + 1] - apos[ia * 3 + 1],
                                 apos[ja * 3 + 2] - apos[ia * 3 + 2]};
            int imin = -1;
            float r2min = INFINITY;
            for (int ipbc = 0; ipbc < npbc; ipbc++) {
                const double x = shifts[ipbc * 3] + d[0], y = shifts[ipbc * 3 + 1] + d[1],
                             z = shifts[ipbc * 3 + 2] + d[2];
                const float r2 = (float)(x * x + y * y + z * z);
                if (r2 < r2min) {
                    r2min = r2;
                    imin = ipbc;
                }
            }
            neighCell[ia * 4 + j] = imin;
        }
}

}  // namespace fc

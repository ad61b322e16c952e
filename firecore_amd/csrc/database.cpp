// Minima database of the exploring mode (host side).
//
// Restates MolecularDatabase::addIfNewDescriptor (MolecularDatabase.h:430-474)
// with the default descriptors of setDescriptors (:377-400): a structure is a
// duplicate of member i when
//   |d - d_i| < 5, d = (sqrt of the inertia tensor's eigenvalues about the
//   centroid, ascending; number of atoms)      (calculatePrincipalAxes :80-140,
//                                                calculateNatoms :142-170)
// and, after superposing its centred positions on the member's centred
// positions by the optimal rotation (computeDistance :527-582 with
// superimposer :723-770),
//   sqrt( sum_i min_{j of i's type} |R x_j - y_i|^2 ) / N < 0.1
// (AtomicConfiguration::dist, AtomicConfiguration.h:19-35: the nearest atom of
// the same type, not a one-to-one assignment).  A duplicate raises the member's
// occurrence count; a new structure becomes a member.
//
// The optimal rotation is computed by Horn's closed form (unit quaternion =
// eigenvector of the largest eigenvalue of the 4x4 symmetric matrix built from
// the correlation sum_k x_k y_k^T, Jacobi eigen-iteration) instead of the
// reference's iterative symmetrisation; both give the rotation that minimises
// sum_k |R x_k - y_k|^2.  Structures of different atom counts are never
// duplicates here (the reference's computeDistance returns -1 for them, which
// the `< 0.1` test of its caller takes as a match; its batches hold replicas
// of one molecule only).
#include "fc_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>

namespace fc {

namespace {

// eigenvalues / eigenvectors of a symmetric n x n matrix (n <= 4) by cyclic
// Jacobi rotations; a is destroyed, eigenvector k is column k of v
void jacobi_sym(int n, double a[4][4], double w[4], double v[4][4]) {
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) v[i][j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 64; sweep++) {
        double off = 0.0;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) off += a[p][q] * a[p][q];
        if (off < 1e-30) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                if (std::fabs(a[p][q]) < 1e-300) continue;
                const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {  // A <- A J
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {  // A <- J^T A
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    const double vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = c * vkp - s * vkq;
                    v[k][q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < n; i++) w[i] = a[i][i];
}

void centroid(int n, const double* x, double c[3]) {
    c[0] = c[1] = c[2] = 0.0;
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) c[k] += x[i * 3 + k];
    for (int k = 0; k < 3; k++) c[k] /= std::max(n, 1);
}

}  // namespace

int db_descriptor(int n, const double* pos, double* d) {
    // calculatePrincipalAxes: unit weights, I = sum (|r|^2 E - r r^T) about the
    // centroid, sqrt of the ascending eigenvalues; skipped below 3 atoms
    int nd = 0;
    if (n >= 3) {
        double c[3];
        centroid(n, pos, c);
        double I[4][4] = {};
        for (int i = 0; i < n; i++) {
            const double r[3] = {pos[i * 3] - c[0], pos[i * 3 + 1] - c[1], pos[i * 3 + 2] - c[2]};
            const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) I[a][b] += (a == b ? r2 : 0.0) - r[a] * r[b];
        }
        double w[4], v[4][4];
        jacobi_sym(3, I, w, v);
        std::sort(w, w + 3);
        for (int k = 0; k < 3; k++) d[nd++] = std::sqrt(std::max(w[k], 0.0));
    }
    d[nd++] = (double)n;  // calculateNatoms
    return nd;
}

double db_distance(int n, const double* x, const int* tx, const double* y, const int* ty) {
    // centre both, rotate x optimally onto y (Horn), then the nearest-same-type sum
    double cx[3], cy[3];
    centroid(n, x, cx);
    centroid(n, y, cy);
    double S[3][3] = {};
    for (int i = 0; i < n; i++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) S[a][b] += (x[i * 3 + a] - cx[a]) * (y[i * 3 + b] - cy[b]);
    double Nm[4][4] = {
        {S[0][0] + S[1][1] + S[2][2], S[1][2] - S[2][1], S[2][0] - S[0][2], S[0][1] - S[1][0]},
        {S[1][2] - S[2][1], S[0][0] - S[1][1] - S[2][2], S[0][1] + S[1][0], S[2][0] + S[0][2]},
        {S[2][0] - S[0][2], S[0][1] + S[1][0], -S[0][0] + S[1][1] - S[2][2], S[1][2] + S[2][1]},
        {S[0][1] - S[1][0], S[2][0] + S[0][2], S[1][2] + S[2][1], -S[0][0] - S[1][1] + S[2][2]}};
    double w[4], v[4][4];
    jacobi_sym(4, Nm, w, v);
    int k = 0;
    for (int i = 1; i < 4; i++)
        if (w[i] > w[k]) k = i;
    double q0 = v[0][k], q1 = v[1][k], q2 = v[2][k], q3 = v[3][k];
    const double qn = std::sqrt(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
    q0 /= qn;
    q1 /= qn;
    q2 /= qn;
    q3 /= qn;
    const double R[3][3] = {
        {q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3, 2 * (q1 * q2 - q0 * q3), 2 * (q1 * q3 + q0 * q2)},
        {2 * (q1 * q2 + q0 * q3), q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3, 2 * (q2 * q3 - q0 * q1)},
        {2 * (q1 * q3 - q0 * q2), 2 * (q2 * q3 + q0 * q1), q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3}};
    std::vector<double> xr((size_t)n * 3);
    for (int i = 0; i < n; i++) {
        const double r[3] = {x[i * 3] - cx[0], x[i * 3 + 1] - cx[1], x[i * 3 + 2] - cx[2]};
        for (int a = 0; a < 3; a++) xr[i * 3 + a] = R[a][0] * r[0] + R[a][1] * r[1] + R[a][2] * r[2];
    }
    double sum = 0.0;
    for (int i = 0; i < n; i++) {  // every atom of y: its nearest x atom of the same type
        const double yi[3] = {y[i * 3] - cy[0], y[i * 3 + 1] - cy[1], y[i * 3 + 2] - cy[2]};
        double best = 1e300;
        for (int j = 0; j < n; j++) {
            if (tx[j] != ty[i]) continue;
            const double d0 = xr[j * 3] - yi[0], d1 = xr[j * 3 + 1] - yi[1], d2 = xr[j * 3 + 2] - yi[2];
            best = std::min(best, d0 * d0 + d1 * d1 + d2 * d2);
        }
        sum += best;
    }
    return std::sqrt(sum);
}

int MinimaDB::add_if_new(int n, const double* pos, const int* types, double E, int isys) {
    total++;
    double d[4];
    const int nd = db_descriptor(n, pos, d);
    for (size_t i = 0; i < members.size(); i++) {
        const Member& m = members[i];
        if (m.natoms != n || m.ndesc != nd) continue;
        double s = 0.0;
        for (int k = 0; k < nd; k++) s += (d[k] - m.desc[k]) * (d[k] - m.desc[k]);
        if (!(std::sqrt(s) < DESC_TOL)) continue;
        if (db_distance(n, pos, types, m.pos.data(), m.types.data()) / n < DIST_TOL) {
            members[i].occurrence++;
            return (int)i;
        }
    }
    Member m;
    m.natoms = n;
    m.pos.assign(pos, pos + (size_t)n * 3);
    m.types.assign(types, types + n);
    m.ndesc = nd;
    for (int k = 0; k < nd; k++) m.desc[k] = d[k];
    m.E = E;
    m.isys = isys;
    m.occurrence = 1;
    members.push_back(std::move(m));
    return -1;
}

}  // namespace fc

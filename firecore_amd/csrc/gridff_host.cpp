// GridFF host side: the NaCl(001) substrate, the node potentials of the
// Pauli / London / Coulomb channels, exact cubic B-spline coefficients and the
// Bspline_PLQd.npy cache file.
//
// Node potentials follow the reference's B-spline grid generation
// (GridFF::evalBsplineRef, GridFF.h:848-890):
//   Pauli   = sum_a sum_img  E_a exp(-2 alpha (r - R0_a))   evalGridFFPoint_Mors
//   London  = sum_a sum_img -2 E_a exp(-alpha (r - R0_a))   (GridFF.h:582-611)
//   Coulomb = sum_a sum_img  k q_a / r                      evalGridFFPoint_Coul
//                                                           (GridFF.h:614-635)
// with E_a = REQ_a.y (sqrt of the well depth), Morse images within 20 A of the
// cell (autoNPBC(cell, nPBC, 20), GridFF.h:43-48, initGridFF :1564) and
// Coulomb images -25..25 in x and y (nPBC_coul, GridFF.h:849).  The reference
// adds 1e-32 to r^2 in the Coulomb term; a node that coincides with an ion
// (the grid origin lies in the top atomic plane) would then hold ~1e17 eV and
// spoil the interpolation far above the surface, so here r^2 is bounded below
// by (0.01 A)^2 -- a change confined to those nodes, whose influence on the
// interpolating spline decays by (2-sqrt 3)^k over k nodes.
// An atom with PLQ = (e^2 E, e E, Q), e = exp(alpha R0) (REQ2PLQ_d,
// Forces.h:611-619) then feels E = PLQ . (Pauli, London, Coulomb).
// The file layout is the reference's: shape (nx,ny,nz,3) float64, C order, z
// fastest (GridFF::pack_Bspline_d, GridFF.h:1203-1217; tryLoad_new :1787-1866).
#include "fc_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <thread>

namespace fc {

static const double COULOMB_K = 14.3996448915;  // eV*A/e^2 (cpp/common/constants.h:9)

// ------------------------------------------------------------------ npy I/O
bool npy_save_f64(const std::string& fname, const std::vector<int>& shape, const double* data) {
    std::string sh = "(";
    size_t n = 1;
    for (size_t i = 0; i < shape.size(); i++) {
        sh += std::to_string(shape[i]);
        sh += (shape.size() == 1 || i + 1 < shape.size()) ? ", " : "";
        n *= (size_t)shape[i];
    }
    sh += ")";
    std::string hdr = "{'descr': '<f8', 'fortran_order': False, 'shape': " + sh + ", }";
    size_t total = 10 + hdr.size() + 1;
    size_t pad = (64 - total % 64) % 64;
    hdr += std::string(pad, ' ');
    hdr += '\n';
    FILE* f = std::fopen(fname.c_str(), "wb");
    if (!f) return false;
    const char magic[8] = {'\x93', 'N', 'U', 'M', 'P', 'Y', 1, 0};
    std::fwrite(magic, 1, 8, f);
    uint16_t hl = (uint16_t)hdr.size();
    std::fwrite(&hl, 2, 1, f);
    std::fwrite(hdr.data(), 1, hdr.size(), f);
    size_t w = std::fwrite(data, sizeof(double), n, f);
    std::fclose(f);
    return w == n;
}

bool npy_load_f64(const std::string& fname, std::vector<int>& shape, std::vector<double>& data) {
    FILE* f = std::fopen(fname.c_str(), "rb");
    if (!f) return false;
    char magic[8];
    if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, "\x93NUMPY", 6) != 0) {
        std::fclose(f);
        return false;
    }
    uint32_t hl = 0;
    if (magic[6] == 1) {
        uint16_t h16;
        if (std::fread(&h16, 2, 1, f) != 1) { std::fclose(f); return false; }
        hl = h16;
    } else {
        if (std::fread(&hl, 4, 1, f) != 1) { std::fclose(f); return false; }
    }
    std::string hdr(hl, ' ');
    if (std::fread(&hdr[0], 1, hl, f) != hl) { std::fclose(f); return false; }
    if (hdr.find("'<f8'") == std::string::npos || hdr.find("'fortran_order': False") == std::string::npos) {
        std::fclose(f);
        return false;
    }
    size_t p0 = hdr.find('(', hdr.find("'shape'"));
    size_t p1 = hdr.find(')', p0);
    shape.clear();
    std::stringstream ss(hdr.substr(p0 + 1, p1 - p0 - 1));
    std::string tok;
    size_t n = 1;
    while (std::getline(ss, tok, ',')) {
        if (tok.find_first_of("0123456789") == std::string::npos) continue;
        int v = std::stoi(tok);
        shape.push_back(v);
        n *= (size_t)v;
    }
    data.resize(n);
    size_t r = std::fread(data.data(), sizeof(double), n, f);
    std::fclose(f);
    return r == n;
}

// ------------------------------------------------------------ substrate
void make_nacl_slab(double a, int nlayers, double q, Surface& s) {
    // conventional square (001) cell, 4 ions per layer, top layer at z=0
    s.apos.clear();
    s.REQs.clear();
    const double RNa = 1.4915, ENa = 0.00130092313, RCl = 1.9735, ECl = 0.00984365166;
    for (int L = 0; L < nlayers; L++) {
        double z = -0.5 * a * L;
        const double xy[4][2] = {{0, 0}, {0.5, 0.5}, {0.5, 0}, {0, 0.5}};
        for (int k = 0; k < 4; k++) {
            bool na = ((k < 2) ^ (L & 1));
            s.apos.push_back(xy[k][0] * a);
            s.apos.push_back(xy[k][1] * a);
            s.apos.push_back(z);
            s.REQs.push_back(na ? RNa : RCl);
            s.REQs.push_back(std::sqrt(na ? ENa : ECl));
            s.REQs.push_back(na ? q : -q);
            s.REQs.push_back(0.0);
        }
    }
    double lv[9] = {a, 0, 0, 0, a, 0, 0, 0, 20.0};
    std::memcpy(s.lvec, lv, sizeof(lv));
}

bool save_surface_xyz(const std::string& fname, const Surface& s) {
    FILE* f = std::fopen(fname.c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "%d\nlvs", s.natoms());
    for (int i = 0; i < 9; i++) std::fprintf(f, " %.10f", s.lvec[i]);
    std::fprintf(f, "\n");
    for (int i = 0; i < s.natoms(); i++) {
        const char* el = (s.REQs[i * 4 + 2] > 0) ? "Na" : "Cl";
        std::fprintf(f, "%s %.10f %.10f %.10f %.6f\n", el, s.apos[i * 3], s.apos[i * 3 + 1], s.apos[i * 3 + 2],
                     s.REQs[i * 4 + 2]);
    }
    std::fclose(f);
    return true;
}

bool load_surface_xyz(const std::string& fname, Surface& s) {
    std::ifstream in(fname);
    if (!in) return false;
    int n = 0;
    std::string line;
    if (!std::getline(in, line)) return false;
    n = std::atoi(line.c_str());
    if (!std::getline(in, line)) return false;
    {
        std::stringstream ss(line);
        std::string tag;
        ss >> tag;
        if (tag != "lvs") return false;
        for (int i = 0; i < 9; i++) ss >> s.lvec[i];
    }
    s.apos.clear();
    s.REQs.clear();
    const double RNa = 1.4915, ENa = 0.00130092313, RCl = 1.9735, ECl = 0.00984365166;
    for (int i = 0; i < n; i++) {
        if (!std::getline(in, line)) return false;
        std::stringstream ss(line);
        std::string el;
        double x, y, z, q = 0;
        ss >> el >> x >> y >> z >> q;
        s.apos.push_back(x);
        s.apos.push_back(y);
        s.apos.push_back(z);
        bool na = (el == "Na");
        s.REQs.push_back(na ? RNa : RCl);
        s.REQs.push_back(std::sqrt(na ? ENa : ECl));
        s.REQs.push_back(q);
        s.REQs.push_back(0.0);
    }
    return true;
}

// ------------------------------------------------------- B-spline fitting
// Solve (c[i-1] + 4 c[i] + c[i+1]) / 6 = v[i] on a periodic line of length n.
static void solve_cyclic_141(int n, const double* v, int stride, double* c) {
    // cyclic Thomas with Sherman-Morrison; a=1/6, b=4/6, c=1/6
    const double a = 1.0 / 6.0, b = 4.0 / 6.0, cc = 1.0 / 6.0;
    std::vector<double> bb(n), x(n), z(n), u(n, 0.0), cp(n), dp(n);
    const double gamma = -b;
    for (int i = 0; i < n; i++) bb[i] = b;
    bb[0] = b - gamma;
    bb[n - 1] = b - a * cc / gamma;
    auto thomas = [&](const std::vector<double>& rhs, std::vector<double>& out) {
        cp[0] = cc / bb[0];
        dp[0] = rhs[0] / bb[0];
        for (int i = 1; i < n; i++) {
            double m = bb[i] - a * cp[i - 1];
            cp[i] = cc / m;
            dp[i] = (rhs[i] - a * dp[i - 1]) / m;
        }
        out[n - 1] = dp[n - 1];
        for (int i = n - 2; i >= 0; i--) out[i] = dp[i] - cp[i] * out[i + 1];
    };
    std::vector<double> rhs(n);
    for (int i = 0; i < n; i++) rhs[i] = v[(size_t)i * stride];
    thomas(rhs, x);
    u[0] = gamma;
    u[n - 1] = cc;
    thomas(u, z);
    double fact = (x[0] + a * x[n - 1] / gamma) / (1.0 + z[0] + a * z[n - 1] / gamma);
    for (int i = 0; i < n; i++) c[(size_t)i * stride] = x[i] - fact * z[i];
}

// Solve the same system on an open line with c[-1]=2c[0]-c[1] and
// c[n]=2c[n-1]-c[n-2] (linear extrapolation), i.e. c[0]=v[0], c[n-1]=v[n-1].
static void solve_open_141(int n, const double* v, int stride, double* c) {
    std::vector<double> cp(n), dp(n);
    const double a = 1.0 / 6.0, b = 4.0 / 6.0, cc = 1.0 / 6.0;
    // rows 0 and n-1 are identity
    cp[0] = 0.0;
    dp[0] = v[0];
    for (int i = 1; i < n - 1; i++) {
        double m = b - a * cp[i - 1];
        cp[i] = cc / m;
        dp[i] = (v[(size_t)i * stride] - a * dp[i - 1]) / m;
    }
    c[(size_t)(n - 1) * stride] = v[(size_t)(n - 1) * stride];
    for (int i = n - 2; i >= 1; i--) c[(size_t)i * stride] = dp[i] - cp[i] * c[(size_t)(i + 1) * stride];
    c[0] = v[0];
}

// Images of the node sums: Morse within 20 A (autoNPBC), Coulomb -25..25.
void gridff_images(const Surface& s, int& nmx, int& nmy, int& ncx, int& ncy) {
    const double la = std::sqrt(s.lvec[0] * s.lvec[0] + s.lvec[1] * s.lvec[1] + s.lvec[2] * s.lvec[2]);
    const double lb = std::sqrt(s.lvec[3] * s.lvec[3] + s.lvec[4] * s.lvec[4] + s.lvec[5] * s.lvec[5]);
    nmx = (int)(20.0 / la);
    nmy = (int)(20.0 / lb);
    ncx = 25;
    ncy = 25;
}

// Node potentials (P, L, Q) of one point, FP64, in the reference's image order
// (atoms outer, images inner; GridFF.h:591-611, 620-634).
void gridff_point(const Surface& s, const double* p, double alpha, int nmx, int nmy, int ncx, int ncy, double* out) {
    const int na = s.natoms();
    const double ax = s.lvec[0], ay = s.lvec[1], bx = s.lvec[3], by = s.lvec[4];
    double P = 0, L = 0, Q = 0;
    for (int a = 0; a < na; a++) {
        const double* A = &s.apos[a * 3];
        const double R0 = s.REQs[a * 4 + 0], E = s.REQs[a * 4 + 1], q = s.REQs[a * 4 + 2];
        for (int jy = -nmy; jy <= nmy; jy++)
            for (int jx = -nmx; jx <= nmx; jx++) {
                const double dx = p[0] - A[0] + jx * ax + jy * bx, dy = p[1] - A[1] + jx * ay + jy * by,
                             dz = p[2] - A[2];
                const double r = std::sqrt(dx * dx + dy * dy + dz * dz + 1e-32);
                const double e = std::exp(-alpha * (r - R0));
                const double eM = e * E;
                P += eM * e;
                L += eM * -2.0;
            }
        for (int jy = -ncy; jy <= ncy; jy++)
            for (int jx = -ncx; jx <= ncx; jx++) {
                const double dx = p[0] - A[0] + jx * ax + jy * bx, dy = p[1] - A[1] + jx * ay + jy * by,
                             dz = p[2] - A[2];
                const double d2 = dx * dx + dy * dy + dz * dz;
                const double r2 = (s.Rdamp > 0.0) ? d2 + s.Rdamp * s.Rdamp : std::max(d2, GRIDFF_R2MIN);
                Q += COULOMB_K * q / std::sqrt(r2);
            }
    }
    out[0] = P;
    out[1] = L;
    out[2] = Q;
}

template <class F>
static void parallel_for(int n, int nthreads, F f);

void gridff_geometry(const Surface& s, int nx, int ny, int nz, double dz, double z0, double alpha, GridFFHost& g) {
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.alphaMorse = alpha;
    const double ax = s.lvec[0], by = s.lvec[4];
    g.step[0] = ax / nx;
    g.step[1] = by / ny;
    g.step[2] = dz;
    // grid origin: cell centred on the substrate origin in x,y
    // (center_cell(-0.5,-0.5,0), MolWorld_sp3_multi.h:2535-2546), z0 = top
    g.pos0[0] = -0.5 * ax;
    g.pos0[1] = -0.5 * by;
    g.pos0[2] = z0;
}

// Node potentials of the whole grid on the host: V (nx,ny,nz,3), z fastest.
void gridff_nodes_host(const Surface& s, const GridFFHost& g, int nthreads, std::vector<double>& V) {
    int nmx, nmy, ncx, ncy;
    gridff_images(s, nmx, nmy, ncx, ncy);
    V.assign((size_t)g.nx * g.ny * g.nz * 3, 0.0);
    parallel_for(g.nx, nthreads, [&](int ix) {
        for (int iy = 0; iy < g.ny; iy++)
            for (int iz = 0; iz < g.nz; iz++) {
                const double p[3] = {g.pos0[0] + ix * g.step[0], g.pos0[1] + iy * g.step[1],
                                     g.pos0[2] + iz * g.step[2]};
                gridff_point(s, p, g.alphaMorse, nmx, nmy, ncx, ncy, &V[(((size_t)ix * g.ny + iy) * g.nz + iz) * 3]);
            }
    });
}

// Exact cubic B-spline interpolation of node values V (nx,ny,nz,3): open in z
// (linear extrapolation at both ends), periodic in y and x; coefficients into
// g.coefs in the same layout.
void gridff_fit(GridFFHost& g, std::vector<double>& V, int nthreads) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const size_t np = (size_t)nx * ny * nz;
    g.coefs.assign(np * 3, 0.0);
    std::vector<double> T(np * 3);
    parallel_for(nx * ny, nthreads, [&](int ixy) {
        for (int ch = 0; ch < 3; ch++)
            solve_open_141(nz, &V[(size_t)ixy * nz * 3 + ch], 3, &T[(size_t)ixy * nz * 3 + ch]);
    });
    parallel_for(nx, nthreads, [&](int ix) {
        for (int iz = 0; iz < nz; iz++)
            for (int ch = 0; ch < 3; ch++) {
                size_t base = ((size_t)ix * ny * nz + iz) * 3 + ch;
                solve_cyclic_141(ny, &T[base], nz * 3, &V[base]);
            }
    });
    parallel_for(ny, nthreads, [&](int iy) {
        for (int iz = 0; iz < nz; iz++)
            for (int ch = 0; ch < 3; ch++) {
                size_t base = ((size_t)iy * nz + iz) * 3 + ch;
                solve_cyclic_141(nx, &V[base], ny * nz * 3, &g.coefs[base]);
            }
    });
}

template <class F>
static void parallel_for(int n, int nthreads, F f) {
    nthreads = std::max(1, std::min(nthreads, n));
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++)
        th.emplace_back([=]() {
            for (int i = t; i < n; i += nthreads) f(i);
        });
    for (auto& x : th) x.join();
}

}  // namespace fc

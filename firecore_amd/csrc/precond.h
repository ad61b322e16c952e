// Rigid-body preconditioner of the FIRE force (device side).
//
// The relaxation of a molecule on a substrate is dominated by soft collective
// motions: sliding/rotation of the whole molecule over the surface and
// rotations of small bonded groups.  FIRE with unit masses moves them at the
// pace of the stiffest bond.  The preconditioned force
//
//   F~_a = F_a + sum_{k : a in S_k} g_k * (t_k + w_k x r_{k,a})
//              + g_c(a) * (t_c + w_c x r_{c,a})
//
// adds, for every bonded star S_k = {node k} + its neighbours and for every
// connected component c, the rigid-body (translation + rotation) projection
// of the force restricted to that group: t = mean force, w = I^-1 L with L
// the torque and I the unit-mass inertia tensor about the group's centroid.
// Each projection is an orthogonal projector in the Euclidean metric, so with
// gains g >= 0 the map F -> F~ is symmetric positive definite: F~ = 0 exactly
// when F = 0, and FIRE on F~ has the same fixed points as on F.  Convergence
// is always judged on the exact force F.  The gains are either constants or
// the outputs of the proposal network (proposal.h).
#pragma once
#include "kernels.h"

namespace fcd {

constexpr int PC_MAXC = 4;  // components per system handled by the molecule-level projection
constexpr int PC_NSUM = 16; // reduced sums per component: n, sum r(3), sum F(3), sum r x F(3), sum r r^T(6)

// Solve (I + eps*tr(I)) w = b for a symmetric 3x3 I = (xx, yy, zz, xy, xz, yz).
// The small relative shift keeps collinear groups (rank-2 inertia) finite: the
// torque of such a group has no component along its axis.
__device__ __forceinline__ V3 solve_sym3(float xx, float yy, float zz, float xy, float xz, float yz, V3 b) {
    const float e = 1e-6f * (xx + yy + zz) + 1e-20f;
    const float a = xx + e, d = yy + e, f = zz + e;
    const float c00 = d * f - yz * yz;
    const float c01 = xz * yz - xy * f;
    const float c02 = xy * yz - d * xz;
    const float c11 = a * f - xz * xz;
    const float c12 = xy * xz - a * yz;
    const float c22 = a * d - xy * xy;
    const float det = a * c00 + xy * c01 + xz * c02;
    const float id = 1.f / det;
    return mk((c00 * b.x + c01 * b.y + c02 * b.z) * id, (c01 * b.x + c11 * b.y + c12 * b.z) * id,
              (c02 * b.x + c12 * b.y + c22 * b.z) * id);
}

// Rigid projection of the force on a star of up to 5 atoms (ok[a] marks the
// members).  r: positions relative to any common origin on entry, relative to
// the centroid on return.  Returns t and w.
__device__ __forceinline__ void star_rigid(const bool* ok, V3* r, const V3* f, V3& t, V3& w) {
    V3 c = zero3(), fs = zero3();
    float n = 0.f;
#pragma unroll
    for (int a = 0; a < 5; a++)
        if (ok[a]) {
            c += r[a];
            fs += f[a];
            n += 1.f;
        }
    const float in = 1.f / n;
    c = c * in;
    t = fs * in;
    V3 L = zero3();
    float xx = 0.f, yy = 0.f, zz = 0.f, xy = 0.f, xz = 0.f, yz = 0.f, q2s = 0.f;
#pragma unroll
    for (int a = 0; a < 5; a++) {
        const V3 q = r[a] - c;
        r[a] = q;
        if (ok[a]) {
            L += cross(q, f[a]);
            const float q2 = dot(q, q);
            q2s += q2;
            xx += q2 - q.x * q.x;
            yy += q2 - q.y * q.y;
            zz += q2 - q.z * q.z;
            xy -= q.x * q.y;
            xz -= q.x * q.z;
            yz -= q.y * q.z;
        }
    }
    // two atoms: the rotations about the bond axis are not rigid motions of
    // the pair, the inertia is singular; the projection onto rotations
    // perpendicular to the axis is w = L / sum |q|^2 exactly
    w = (n == 2.f) ? L * (1.f / q2s) : solve_sym3(xx, yy, zz, xy, xz, yz, L);
}

// Molecule-level projection from the reduced sums S[0..15] of one component:
// n, sum r, sum F, sum r x F, sum r r^T (xx, yy, zz, xy, xz, yz); r relative to
// a common origin.  Returns the centroid c, the mean force t and w.
__device__ __forceinline__ void comp_rigid(const float* S, V3& c, V3& t, V3& w) {
    const float n = fmaxf(S[0], 1.f);
    const float in = 1.f / n;
    c = mk(S[1], S[2], S[3]) * in;
    const V3 fs = mk(S[4], S[5], S[6]);
    t = fs * in;
    // torque about the centroid: sum (r - c) x F = sum r x F - c x sum F
    const V3 L = mk(S[7], S[8], S[9]) - cross(c, fs);
    // second moments about the centroid: sum (r-c)(r-c)^T = sum r r^T - n c c^T
    const float mxx = S[10] - n * c.x * c.x, myy = S[11] - n * c.y * c.y, mzz = S[12] - n * c.z * c.z;
    const float mxy = S[13] - n * c.x * c.y, mxz = S[14] - n * c.x * c.z, myz = S[15] - n * c.y * c.z;
    const float tr = mxx + myy + mzz;
    w = (S[0] == 2.f) ? L * (1.f / tr) : solve_sym3(tr - mxx, tr - myy, tr - mzz, -mxy, -mxz, -myz, L);
}

// Wave-wide reductions (64 lanes).
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
// Wave totals of 4 per-lane values by a reduce-scatter butterfly (7
// shuffles instead of 4 full reductions of 6): lane l returns the total of
// value l >> 4.
__device__ __forceinline__ float wave_sum4(float v0, float v1, float v2, float v3) {
    const int lane = threadIdx.x & 63;
    const bool u32 = (lane & 32) != 0;
    const float a = (u32 ? v2 : v0) + __shfl_xor(u32 ? v0 : v2, 32, 64);
    const float b = (u32 ? v3 : v1) + __shfl_xor(u32 ? v1 : v3, 32, 64);
    const bool u16 = (lane & 16) != 0;
    float s = (u16 ? b : a) + __shfl_xor(u16 ? a : b, 16, 64);
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    return s;
}
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Preconditioned force of the atoms of one system (pi-vectors keep their
// exact force).  Entry: fv0 = exact assembled force of atom tid (tid < N);
// every thread of the block calls this.  s_fnb receives the exact atom forces;
// s_fself / s_fng (whose recoils have been consumed) hold the star terms.
// Star gain g_s and molecule gain g_c are P.pc_gstar / P.pc_grigid, or the
// learned gains in s_gain (per node) and s_gain[M + c] (per component).
template <int BLOCK>
__device__ __forceinline__ void precondition(const RunParams& P, const float4* s_pos, float4* s_fnb, float4* s_fself,
                                             float4* s_fng, float* s_cred, float* s_csum, const float* s_gain,
                                             bool learned, int N, int M, int nc, int myc, bool hasAtom,
                                             bool hasNode, int4 ng, int4 bk0, V3& fv0) {
    constexpr int NW = BLOCK / 64;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wid = tid >> 6;
    const V3 org = xyz(s_pos[0]);
    V3 r = zero3();
    if (hasAtom) {
        r = xyz(s_pos[tid]) - org;
        s_fnb[tid] = make_float4(fv0.x, fv0.y, fv0.z, 0.f);
    }
    // per-component sums: n, r, F, r x F, r r^T; wave totals into s_cred
    {
        const V3 f = fv0;
        const V3 rf = cross(r, f);
        for (int c = 0; c < nc; c++) {
            const float m = (hasAtom && myc == c) ? 1.f : 0.f;
            float* dst = s_cred + (wid * PC_MAXC + c) * PC_NSUM + (lane >> 4);
            const float t0 = wave_sum4(m, m * r.x, m * r.y, m * r.z);
            if ((lane & 15) == 0) dst[0] = t0;
            const float t1 = wave_sum4(m * f.x, m * f.y, m * f.z, m * rf.x);
            if ((lane & 15) == 0) dst[4] = t1;
            const float t2 = wave_sum4(m * rf.y, m * rf.z, m * r.x * r.x, m * r.y * r.y);
            if ((lane & 15) == 0) dst[8] = t2;
            const float t3 = wave_sum4(m * r.z * r.z, m * r.x * r.y, m * r.x * r.z, m * r.y * r.z);
            if ((lane & 15) == 0) dst[12] = t3;
        }
    }
    __syncthreads();
    // one thread per component: block totals and the rigid-body fit, kept in
    // s_csum[c*16]: count, centroid (1..3), mean force t (4..6), angular rate w (7..9)
    if (tid < nc) {
        float S[PC_NSUM];
#pragma unroll
        for (int k = 0; k < PC_NSUM; k++) {
            float v = 0.f;
#pragma unroll
            for (int w = 0; w < NW; w++) v += s_cred[(w * PC_MAXC + tid) * PC_NSUM + k];
            S[k] = v;
        }
        V3 cen, t, w;
        comp_rigid(S, cen, t, w);
        float* o = s_csum + tid * PC_NSUM;
        o[0] = S[0];
        o[1] = cen.x;
        o[2] = cen.y;
        o[3] = cen.z;
        o[4] = t.x;
        o[5] = t.y;
        o[6] = t.z;
        o[7] = w.x;
        o[8] = w.y;
        o[9] = w.z;
    }
    if (hasNode) {
        // star of node tid: the node (slot 0) and its bonded neighbours (slots 1..4, masked when -1)
        V3 rs[5], fs[5];
        bool ok[5];
        const int nb4[4] = {ng.x, ng.y, ng.z, ng.w};
        const V3 pk = xyz(s_pos[tid]);
        rs[0] = zero3();
        fs[0] = xyz(s_fnb[tid]);
        ok[0] = true;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            ok[j + 1] = nb4[j] >= 0;
            rs[j + 1] = ok[j + 1] ? xyz(s_pos[nb4[j]]) - pk : zero3();
            fs[j + 1] = ok[j + 1] ? xyz(s_fnb[nb4[j]]) : zero3();
        }
        V3 t, w;
        star_rigid(ok, rs, fs, t, w);
        const float g = learned ? s_gain[tid] : P.pc_gstar;
        const V3 d0 = (t + cross(w, rs[0])) * g;
        s_fself[tid] = make_float4(d0.x, d0.y, d0.z, 0.f);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const V3 d = (t + cross(w, rs[j + 1])) * g;
            s_fng[tid * 4 + j] = make_float4(d.x, d.y, d.z, 0.f);
        }
    }
    __syncthreads();
    if (hasAtom) {
        V3 f = fv0;
        if (tid < M) f += xyz(s_fself[tid]);
        const int b4[4] = {bk0.x, bk0.y, bk0.z, bk0.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (b4[k] < 0) break;
            f += xyz(s_fng[b4[k]]);
        }
        const float* o = s_csum + myc * PC_NSUM;
        const V3 cen = mk(o[1], o[2], o[3]), t = mk(o[4], o[5], o[6]), w = mk(o[7], o[8], o[9]);
        const float gc = learned ? s_gain[M + myc] : P.pc_grigid;
        f += (t + cross(w, r - cen)) * gc;
        fv0 = f;
    }
}

}  // namespace fcd

// Training of the proposal network on the device.
//
// One training step over a batch of relaxation snapshots (every system of the
// batch holds one snapshot geometry x0 with its exact forces F0 and energy E0):
//   1. k_train_fwd       features + forward pass (activations kept for the
//                        backward pass), gains g(theta)
//   2. k_precond         preconditioned force F~(g) (split_kernels.hip)
//   3. k_train_displace  x' = x0 + eta * F~ (per-atom step capped at dmax)
//   4. split kernels     exact forces F' and energy E' at x'
//   5. k_train_loss_grad loss_s = (E'_s - E0_s) / N_s and dloss/d(outputs):
//                        dE'/dg_k = -eta F'.P_k(F0) for star k, the same with
//                        the component projection for the molecule gain
//   6. k_train_bwd       backward pass, dense layers on the matrix cores
//                        (bf16 operands, FP32 accumulation); per-system
//                        parameter gradients
//   7. k_grad_reduce     mean over systems in a fixed order (deterministic)
//   8. k_adam            Adam update of the FP32 master parameters, bf16 copy
//                        of the dense weights for the matrix cores
// The loss is the mean per-atom energy change of one preconditioned steepest-
// descent step from the snapshot: gains that lower the energy more are better.
#include "kernels.h"
#include "precond.h"
#include "proposal.h"

namespace fcd {

constexpr int NP_W1 = 64 * 16, NP_B1 = 64, NP_W2 = 64 * 128, NP_B2 = 64, NP_W3 = 64 * 128, NP_B3 = 64, NP_W4 = 128,
              NP_B4 = 2;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NP_W1, OFF_W2 = OFF_B1 + NP_B1, OFF_B2 = OFF_W2 + NP_W2,
              OFF_B3 = OFF_B2 + NP_B2 + NP_W3, OFF_W3 = OFF_B2 + NP_B2, OFF_W4 = OFF_B3 + NP_B3,
              OFF_B4 = OFF_W4 + NP_W4, NPARAM = OFF_B4 + NP_B4;

struct TrainBufs {
    const float4* x0;        // [sum V] snapshot positions
    const float4* F0;        // [sum V] exact forces at x0
    const double* E0;        // [nsys]
    unsigned short* X;       // [sum N][16] features (bf16)
    float* z;                // [3 * 64 * sum N] pre-activations, per system [3][N][64]
    unsigned short* hs;      // [3 * 64 * sum N] activations (bf16), per system [3][N][64]
    float* out;              // [sum N][2] network outputs
    float* dout;             // [sum N][2] dloss/doutputs
    float* gnode;            // [sum M] star gains
    float* gcomp;            // [nsys * PC_MAXC] molecule gains
    const float4* pcforce;   // [sum V] preconditioned force
    float* loss;             // [nsys]
    float* gpart;            // [nsys][NPARAM]
    float eta, dmax;
};

// ---------------------------------------------------------------- 1. forward
__global__ void __launch_bounds__(256) k_train_fwd(BatchDev B, RunParams P, TrainBufs T) {
    constexpr int BLOCK = 256;
    __shared__ float4 s_pos[2 * BLOCK];
    __shared__ float4 s_F[BLOCK];
    __shared__ __align__(16) unsigned short s_h[BLOCK * MLP_HS];
    __shared__ float s_out[2 * BLOCK];
    const int s = blockIdx.x;
    const SysInfo I = B.info[s];
    const int N = I.natoms, M = I.nnode;
    if (N > BLOCK || M > BLOCK) return;
    const int tid = threadIdx.x;
    for (int v = tid; v < N + M; v += BLOCK) s_pos[v] = T.x0[I.v0 + v];
    if (tid < N) s_F[tid] = T.F0[I.v0 + tid];
    __syncthreads();
    if (tid < N) {
        unsigned short* hrow = s_h + tid * MLP_HS;
        mlp_feature_row(P, s_pos, s_F, B.neighs + I.a0, B.elem + I.a0, N, M, tid, B.PLQs[I.a0 + tid], hrow);
#pragma unroll
        for (int k = 0; k < MLP_NIN; k++) T.X[(size_t)(I.a0 + tid) * MLP_NIN + k] = hrow[k];
    }
    __syncthreads();
    const size_t zo = (size_t)3 * MLP_NH * I.a0;
    mlp_forward<BLOCK>(P.mlp, s_h, B.neighs + I.a0, N, s_out, T.z + zo, T.hs + zo);
    if (tid < N) {
        T.out[(I.a0 + tid) * 2] = s_out[tid * 2];
        T.out[(I.a0 + tid) * 2 + 1] = s_out[tid * 2 + 1];
    }
    if (tid < M) T.gnode[I.n0 + tid] = P.pc_gstar * sigmoidf_(s_out[tid * 2]);
    if (tid < PC_MAXC) {
        float sum = 0.f;
        int cnt = 0;
        for (int i = 0; i < N; i++)
            if (B.comp[I.a0 + i] == tid) {
                sum += s_out[i * 2 + 1];
                cnt++;
            }
        T.gcomp[s * PC_MAXC + tid] = cnt ? P.pc_grigid * sigmoidf_(sum / (float)cnt) : 0.f;
    }
}

// ---------------------------------------------------------------- 3. displace
__global__ void __launch_bounds__(256) k_train_displace(BatchDev B, TrainBufs T) {
    const int s = blockIdx.x;
    const SysInfo I = B.info[s];
    const int N = I.natoms, V = I.natoms + I.nnode;
    for (int v = threadIdx.x; v < V; v += blockDim.x) {
        float4 x = T.x0[I.v0 + v];
        if (v < N) {
            const float4 f = T.pcforce[I.v0 + v];
            V3 d = mk(f.x, f.y, f.z) * T.eta;
            const float dl = sqrtf(dot(d, d));
            if (dl > T.dmax) d = d * (T.dmax / dl);
            x.x += d.x;
            x.y += d.y;
            x.z += d.z;
        }
        B.atoms[I.v0 + v] = x;
    }
}

// ---------------------------------------------------------------- 5. loss gradient
__global__ void __launch_bounds__(256) k_train_loss_grad(BatchDev B, RunParams P, TrainBufs T) {
    constexpr int BLOCK = 256;
    constexpr int NW = BLOCK / 64;
    __shared__ float4 s_pos[BLOCK];
    __shared__ float4 s_F0[BLOCK];
    __shared__ float4 s_F1[BLOCK];
    __shared__ float s_red[NW][PC_MAXC][PC_NSUM + 1];
    __shared__ float s_sum[PC_MAXC][PC_NSUM + 1];
    const int s = blockIdx.x;
    const SysInfo I = B.info[s];
    const int N = I.natoms, M = I.nnode;
    if (N > BLOCK || M > BLOCK) return;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const bool hasAtom = tid < N;
    if (hasAtom) {
        s_pos[tid] = T.x0[I.v0 + tid];
        s_F0[tid] = T.F0[I.v0 + tid];
        s_F1[tid] = B.aforce[I.v0 + tid];
    }
    __syncthreads();
    const float scale = -T.eta / (float)N;
    // star gains: dL/dg_k = -eta/N * sum_{a in S_k} F'_a . P_k(F0)_a
    if (tid < M) {
        const int4 ng = B.neighs[I.a0 + tid];
        const int nb4[4] = {ng.x, ng.y, ng.z, ng.w};
        V3 rs[5], fs[5];
        bool ok[5];
        const V3 pk = xyz(s_pos[tid]);
        rs[0] = zero3();
        fs[0] = xyz(s_F0[tid]);
        ok[0] = true;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            ok[j + 1] = nb4[j] >= 0;
            rs[j + 1] = ok[j + 1] ? xyz(s_pos[nb4[j]]) - pk : zero3();
            fs[j + 1] = ok[j + 1] ? xyz(s_F0[nb4[j]]) : zero3();
        }
        V3 t, w;
        star_rigid(ok, rs, fs, t, w);
        float g = dot(xyz(s_F1[tid]), t + cross(w, rs[0]));
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (ok[j + 1]) g += dot(xyz(s_F1[nb4[j]]), t + cross(w, rs[j + 1]));
        const float o = T.out[(I.a0 + tid) * 2];
        const float sg = sigmoidf_(o);
        T.dout[(I.a0 + tid) * 2] = scale * g * P.pc_gstar * sg * (1.f - sg);
    } else if (hasAtom) {
        T.dout[(I.a0 + tid) * 2] = 0.f;
    }
    // molecule gains: dL/dg_c = -eta/N * sum_{a in c} F'_a . (t_c + w_c x (r_a - c_c)),
    // with (c_c, t_c, w_c) from the component sums of x0 and F0 as in precondition()
    const int nc = B.ncomp[s];
    const int myc = hasAtom ? B.comp[I.a0 + tid] : 0;
    const V3 org = xyz(s_pos[0]);
    const V3 r = hasAtom ? xyz(s_pos[tid]) - org : zero3();
    const V3 f0 = hasAtom ? xyz(s_F0[tid]) : zero3();
    const V3 f1 = hasAtom ? xyz(s_F1[tid]) : zero3();
    const V3 rf0 = cross(r, f0);
    for (int c = 0; c < nc; c++) {
        const bool mine = hasAtom && myc == c;
        const float val[PC_NSUM] = {1.f, r.x, r.y, r.z, f0.x, f0.y, f0.z, rf0.x, rf0.y, rf0.z,
                                    r.x * r.x, r.y * r.y, r.z * r.z, r.x * r.y, r.x * r.z, r.y * r.z};
#pragma unroll
        for (int k = 0; k < PC_NSUM; k++) {
            const float v = wave_sum(mine ? val[k] : 0.f);
            if (lane == 0) s_red[wid][c][k] = v;
        }
    }
    __syncthreads();
    if (tid < nc * PC_NSUM) {
        const int c = tid / PC_NSUM, k = tid % PC_NSUM;
        float v = 0.f;
        for (int w2 = 0; w2 < NW; w2++) v += s_red[w2][c][k];
        s_sum[c][k] = v;
    }
    __syncthreads();
    // per-atom contribution F'_a . (t + w x (r_a - c))
    float contrib = 0.f;
    if (hasAtom) {
        V3 cen, t, w;
        comp_rigid(&s_sum[myc][0], cen, t, w);
        contrib = dot(f1, t + cross(w, r - cen));
    }
    for (int c = 0; c < nc; c++) {
        const float v = wave_sum((hasAtom && myc == c) ? contrib : 0.f);
        if (lane == 0) s_red[wid][c][PC_NSUM] = v;
    }
    __syncthreads();
    if (tid < nc) {
        float v = 0.f;
        for (int w2 = 0; w2 < NW; w2++) v += s_red[w2][tid][PC_NSUM];
        s_sum[tid][PC_NSUM] = v;
    }
    __syncthreads();
    if (hasAtom) {
        const int c = myc;
        const float ncnt = fmaxf(s_sum[c][0], 1.f);
        // mean of o1 over the component, as in k_train_fwd
        float sum = 0.f;
        for (int i = 0; i < N; i++)
            if (B.comp[I.a0 + i] == c) sum += T.out[(I.a0 + i) * 2 + 1];
        const float sg = sigmoidf_(sum / ncnt);
        T.dout[(I.a0 + tid) * 2 + 1] = scale * s_sum[c][PC_NSUM] * P.pc_grigid * sg * (1.f - sg) / ncnt;
    }
    if (tid == 0) T.loss[s] = (float)((B.Esys[s] - T.E0[s]) / (double)N);
}

// ---------------------------------------------------------------- 6. backward
// dW (64 x KIN) += dZ^T A over the atoms of one system.  A rows: own
// activation row (chunks < KS) or the bonded-neighbour sum (chunks >= KS).
template <int KIN>
__device__ __forceinline__ void bwd_weight_grad(const unsigned short* dZ, const unsigned short* hin, const int4* nb,
                                                int n, float* gW) {
    constexpr int KS = (KIN == 16) ? 1 : 4;
    constexpr int NKB = KIN / 16;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nrb = (n + 15) >> 4;
    // 4 x NKB output tiles over 4 waves
    for (int t = wid; t < 4 * NKB; t += 4) {
        const int ob = t / NKB, kb = t % NKB;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kc = 0; kc < nrb; kc++) {
            bf16x4 a, b;
            const int o = ob * 16 + (lane & 15);
            const int col = (kb % KS) * 16 + (lane & 15);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int atom = kc * 16 + 4 * (lane >> 4) + i;
                a[i] = (short)((atom < n) ? dZ[atom * MLP_HS + o] : 0);
                unsigned short bv = 0;
                if (atom < n) {
                    if (kb < KS) {
                        bv = hin[atom * MLP_HS + col];
                    } else {
                        const int4 n4 = nb[atom];
                        const int ns[4] = {n4.x, n4.y, n4.z, n4.w};
                        float sacc = 0.f;
#pragma unroll
                        for (int q = 0; q < 4; q++)
                            if (ns[q] >= 0) sacc += bf2f(hin[ns[q] * MLP_HS + col]);
                        bv = f2bf(sacc);
                    }
                }
                b[i] = (short)bv;
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
        // D[o][k]: lane holds rows 4*(lane/16)+i of column lane%16
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int o = ob * 16 + 4 * (lane >> 4) + i;
            const int k = kb * 16 + (lane & 15);
            gW[o * KIN + k] = acc[i];
        }
    }
}

// dh (n x 64) = [dZ_i ; sum_nb dZ_j] [W_self^T ; W_nb^T]^T, i.e. the gradient
// through a 128-input layer back to its 64 inputs (self and neighbour sums
// share the input rows).  Result tiles in acc (at most 16 per wave).
__device__ __forceinline__ void bwd_input_grad(const unsigned short* dZ, const int4* nb, const unsigned short* W,
                                               int n, f32x4* acc) {
    constexpr int NW = 4;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nrb = (n + 15) >> 4;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
        const int t = wid + q * NW;
        if (t < nrb * 4) {
            const int rb = t >> 2, cb = t & 3;
#pragma unroll
            for (int kc = 0; kc < 8; kc++) {
                const bf16x4 a = mlp_a_frag<4>(dZ, nb, n, rb, kc, lane);
                bf16x4 b;
                const int col = cb * 16 + (lane & 15) + ((kc >= 4) ? 64 : 0);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int o = (kc & 3) * 16 + 4 * (lane >> 4) + i;
                    b[i] = (short)W[o * 128 + col];
                }
                acc[q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[q], 0, 0, 0);
            }
        }
    }
}

// dz = dh * silu'(z) written as bf16 rows into dZ (after the caller's barrier)
__device__ __forceinline__ void bwd_store_dz(unsigned short* dZ, const float* z, int n, const f32x4* acc) {
    constexpr int NW = 4;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nrb = (n + 15) >> 4;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int t = wid + q * NW;
        if (t < nrb * 4) {
            const int rb = t >> 2, cb = t & 3;
            const int col = cb * 16 + (lane & 15);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int row = rb * 16 + 4 * (lane >> 4) + i;
                if (row < n) {
                    const float zz = z[row * MLP_NH + col];
                    const float sg = 1.f / (1.f + __expf(-zz));
                    dZ[row * MLP_HS + col] = f2bf(acc[q][i] * sg * (1.f + zz * (1.f - sg)));
                }
            }
        }
    }
}

__device__ __forceinline__ void bwd_bias_grad(const unsigned short* dZ, int n, float* gb) {
    for (int k = threadIdx.x; k < MLP_NH; k += blockDim.x) {
        float s = 0.f;
        for (int i = 0; i < n; i++) s += bf2f(dZ[i * MLP_HS + k]);
        gb[k] = s;
    }
}

__global__ void __launch_bounds__(256) k_train_bwd(BatchDev B, RunParams P, TrainBufs T) {
    constexpr int BLOCK = 256;
    __shared__ __align__(16) unsigned short s_h[BLOCK * MLP_HS];
    __shared__ __align__(16) unsigned short s_dz[BLOCK * MLP_HS];
    __shared__ float s_do[2 * BLOCK];
    const int s = blockIdx.x;
    const SysInfo I = B.info[s];
    const int n = I.natoms;
    if (n > BLOCK || I.nnode > BLOCK) return;
    const int tid = threadIdx.x;
    const int4* nb = B.neighs + I.a0;
    const size_t zo = (size_t)3 * MLP_NH * I.a0;
    const float* z1 = T.z + zo;
    const float* z2 = z1 + n * MLP_NH;
    const float* z3 = z2 + n * MLP_NH;
    const unsigned short* h1 = T.hs + zo;
    const unsigned short* h2 = h1 + n * MLP_NH;
    const unsigned short* h3 = h2 + n * MLP_NH;
    float* G = T.gpart + (size_t)s * NPARAM;
    if (tid < n) {
        s_do[tid * 2] = T.dout[(I.a0 + tid) * 2];
        s_do[tid * 2 + 1] = T.dout[(I.a0 + tid) * 2 + 1];
    }
    for (int e = tid; e < n * MLP_NH; e += BLOCK) s_h[(e >> 6) * MLP_HS + (e & 63)] = h3[e];
    __syncthreads();
    // head: dW4[j][k] = sum_i do_ij h3_ik, db4 = sum_i do_i
    if (tid < 2 * MLP_NH) {
        const int j = tid >> 6, k = tid & 63;
        float sacc = 0.f;
        for (int i = 0; i < n; i++) sacc += s_do[i * 2 + j] * bf2f(s_h[i * MLP_HS + k]);
        G[OFF_W4 + j * MLP_NH + k] = sacc;
    }
    if (tid < 2) {
        float sacc = 0.f;
        for (int i = 0; i < n; i++) sacc += s_do[i * 2 + tid];
        G[OFF_B4 + tid] = sacc;
    }
    // dz3 = (do W4) * silu'(z3)
    for (int e = tid; e < n * MLP_NH; e += BLOCK) {
        const int i = e >> 6, k = e & 63;
        const float dh = s_do[i * 2] * P.mlp.W4[k] + s_do[i * 2 + 1] * P.mlp.W4[MLP_NH + k];
        const float zz = z3[e];
        const float sg = 1.f / (1.f + __expf(-zz));
        s_dz[i * MLP_HS + k] = f2bf(dh * sg * (1.f + zz * (1.f - sg)));
    }
    __syncthreads();
    bwd_bias_grad(s_dz, n, G + OFF_B3);
    // layer 3: input rows h2
    for (int e = tid; e < n * MLP_NH; e += BLOCK) s_h[(e >> 6) * MLP_HS + (e & 63)] = h2[e];
    __syncthreads();
    bwd_weight_grad<128>(s_dz, s_h, nb, n, G + OFF_W3);
    f32x4 acc[16];
    bwd_input_grad(s_dz, nb, P.mlp.W3, n, acc);
    __syncthreads();
    bwd_store_dz(s_dz, z2, n, acc);
    __syncthreads();
    bwd_bias_grad(s_dz, n, G + OFF_B2);
    // layer 2: input rows h1
    for (int e = tid; e < n * MLP_NH; e += BLOCK) s_h[(e >> 6) * MLP_HS + (e & 63)] = h1[e];
    __syncthreads();
    bwd_weight_grad<128>(s_dz, s_h, nb, n, G + OFF_W2);
    bwd_input_grad(s_dz, nb, P.mlp.W2, n, acc);
    __syncthreads();
    bwd_store_dz(s_dz, z1, n, acc);
    __syncthreads();
    bwd_bias_grad(s_dz, n, G + OFF_B1);
    // layer 1: input rows = features
    for (int e = tid; e < n * MLP_NIN; e += BLOCK)
        s_h[(e >> 4) * MLP_HS + (e & 15)] = T.X[(size_t)I.a0 * MLP_NIN + e];
    __syncthreads();
    bwd_weight_grad<16>(s_dz, s_h, nb, n, G + OFF_W1);
}

// ---------------------------------------------------------------- 7. reduce, 8. Adam
__global__ void k_grad_reduce(const float* gpart, int nsys, float* grad, const float* loss, float* loss_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < NPARAM) {
        float s = 0.f;
        for (int k = 0; k < nsys; k++) s += gpart[(size_t)k * NPARAM + p];
        grad[p] = s / (float)nsys;
    }
    if (p == 0) {
        double l = 0.0;
        for (int k = 0; k < nsys; k++) l += loss[k];
        loss_out[0] = (float)(l / nsys);
    }
}

__device__ __forceinline__ unsigned short f2bf_d(float f) { return f2bf(f); }

__global__ void k_adam(float* theta, float* m, float* v, const float* grad, float lr, float b1, float b2, float eps,
                       float bc1, float bc2, float clip, unsigned short* Wbf, float* Ff) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    float g = grad[p];
    g = fminf(fmaxf(g, -clip), clip);
    const float mm = b1 * m[p] + (1.f - b1) * g;
    const float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float th = theta[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    theta[p] = th;
    // device copies: bf16 W1 | W2 | W3, FP32 b1 | b2 | b3 | W4 | b4
    if (p < OFF_B1) Wbf[p] = f2bf_d(th);
    else if (p < OFF_W2) Ff[p - OFF_B1] = th;
    else if (p < OFF_B2) Wbf[NP_W1 + (p - OFF_W2)] = f2bf_d(th);
    else if (p < OFF_W3) Ff[64 + (p - OFF_B2)] = th;
    else if (p < OFF_B3) Wbf[NP_W1 + NP_W2 + (p - OFF_W3)] = f2bf_d(th);
    else if (p < OFF_W4) Ff[128 + (p - OFF_B3)] = th;
    else Ff[192 + (p - OFF_W4)] = th;
}

}  // namespace fcd

extern "C" int fc_mlp_nparam() { return fcd::NPARAM; }

extern "C" hipError_t fc_launch_train_fwd(hipStream_t st, const fcd::BatchDev* B, const fcd::RunParams* P,
                                          const fcd::TrainBufs* T) {
    hipLaunchKernelGGL(fcd::k_train_fwd, dim3(B->nsys), dim3(256), 0, st, *B, *P, *T);
    return hipGetLastError();
}
extern "C" hipError_t fc_launch_train_displace(hipStream_t st, const fcd::BatchDev* B, const fcd::TrainBufs* T) {
    hipLaunchKernelGGL(fcd::k_train_displace, dim3(B->nsys), dim3(256), 0, st, *B, *T);
    return hipGetLastError();
}
extern "C" hipError_t fc_launch_train_loss_grad(hipStream_t st, const fcd::BatchDev* B, const fcd::RunParams* P,
                                                const fcd::TrainBufs* T) {
    hipLaunchKernelGGL(fcd::k_train_loss_grad, dim3(B->nsys), dim3(256), 0, st, *B, *P, *T);
    return hipGetLastError();
}
extern "C" hipError_t fc_launch_train_bwd(hipStream_t st, const fcd::BatchDev* B, const fcd::RunParams* P,
                                          const fcd::TrainBufs* T) {
    hipLaunchKernelGGL(fcd::k_train_bwd, dim3(B->nsys), dim3(256), 0, st, *B, *P, *T);
    return hipGetLastError();
}
extern "C" hipError_t fc_launch_grad_reduce(hipStream_t st, const float* gpart, int nsys, float* grad,
                                            const float* loss, float* loss_out) {
    hipLaunchKernelGGL(fcd::k_grad_reduce, dim3((fcd::NPARAM + 255) / 256), dim3(256), 0, st, gpart, nsys, grad,
                       loss, loss_out);
    return hipGetLastError();
}
extern "C" hipError_t fc_launch_adam(hipStream_t st, float* theta, float* m, float* v, const float* grad, float lr,
                                     float b1, float b2, float eps, float bc1, float bc2, float clip,
                                     unsigned short* Wbf, float* Ff) {
    hipLaunchKernelGGL(fcd::k_adam, dim3((fcd::NPARAM + 255) / 256), dim3(256), 0, st, theta, m, v, grad, lr, b1, b2,
                       eps, bc1, bc2, clip, Wbf, Ff);
    return hipGetLastError();
}

namespace fcd {
// Gains of every running system from the network at its current positions
// (inference between relaxation chunks).  F = the exact forces written back
// by the relaxation (those of the last evaluated step).
__global__ void __launch_bounds__(256) k_mlp_gains(BatchDev B, RunParams P, float* gnode, float* gcomp) {
    constexpr int BLOCK = 256;
    __shared__ float4 s_pos[2 * BLOCK];
    __shared__ float4 s_F[BLOCK];
    __shared__ __align__(16) unsigned short s_h[BLOCK * MLP_HS];
    __shared__ float s_out[2 * BLOCK];
    const int s = blockIdx.x;
    if (B.status[s] != ST_RUNNING) return;
    const SysInfo I = B.info[s];
    const int N = I.natoms, M = I.nnode;
    if (N > BLOCK || M > BLOCK) return;
    const int tid = threadIdx.x;
    for (int v = tid; v < N + M; v += BLOCK) s_pos[v] = B.atoms[I.v0 + v];
    if (tid < N) s_F[tid] = B.aforce[I.v0 + tid];
    __syncthreads();
    if (tid < N)
        mlp_feature_row(P, s_pos, s_F, B.neighs + I.a0, B.elem + I.a0, N, M, tid, B.PLQs[I.a0 + tid],
                        s_h + tid * MLP_HS);
    __syncthreads();
    mlp_forward<BLOCK>(P.mlp, s_h, B.neighs + I.a0, N, s_out, nullptr, nullptr);
    if (tid < M) gnode[I.n0 + tid] = P.pc_gstar * sigmoidf_(s_out[tid * 2]);
    if (tid < PC_MAXC) {
        float sum = 0.f;
        int cnt = 0;
        for (int i = 0; i < N; i++)
            if (B.comp[I.a0 + i] == tid) {
                sum += s_out[i * 2 + 1];
                cnt++;
            }
        gcomp[s * PC_MAXC + tid] = cnt ? P.pc_grigid * sigmoidf_(sum / (float)cnt) : 0.f;
    }
}

// Queue counters of the persistent kernel back to zero (a kernel, so that the
// chunk loop can be captured in a graph without memset nodes).
__global__ void k_zero_int(int* a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

// Number of systems still running, into *out (systems with skip[s] != 0 not
// counted; skip may be null).
__global__ void k_count_running(const int* status, int nsys, int* out, const int* skip) {
    __shared__ int cnt;
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    int c = 0;
    for (int s = threadIdx.x; s < nsys; s += blockDim.x)
        c += (status[s] == ST_RUNNING && !(skip && skip[s])) ? 1 : 0;
    atomicAdd(&cnt, c);
    __syncthreads();
    if (threadIdx.x == 0) out[0] = cnt;
}

__global__ void __launch_bounds__(256) k_gather_snap(BatchDev B, const float4* snap, const int* idx) {
    const int s = blockIdx.x;
    const SysInfo I = B.info[s];
    const float4* src = snap + (size_t)idx[s] * B.ntotV + I.v0;
    for (int v = threadIdx.x; v < I.natoms + I.nnode; v += blockDim.x) B.atoms[I.v0 + v] = src[v];
}
}  // namespace fcd

extern "C" hipError_t fc_launch_gather_snap(hipStream_t st, const fcd::BatchDev* B, const float4* snap,
                                            const int* idx) {
    hipLaunchKernelGGL(fcd::k_gather_snap, dim3(B->nsys), dim3(256), 0, st, *B, snap, idx);
    return hipGetLastError();
}

extern "C" hipError_t fc_launch_mlp_gains(hipStream_t st, const fcd::BatchDev* B, const fcd::RunParams* P, float* gnode,
                                          float* gcomp) {
    hipLaunchKernelGGL(fcd::k_mlp_gains, dim3(B->nsys), dim3(256), 0, st, *B, *P, gnode, gcomp);
    return hipGetLastError();
}
extern "C" hipError_t fc_launch_zero_int(hipStream_t st, int* a, int n) {
    hipLaunchKernelGGL(fcd::k_zero_int, dim3((n + 255) / 256), dim3(256), 0, st, a, n);
    return hipGetLastError();
}
extern "C" hipError_t fc_launch_count_running(hipStream_t st, const int* status, int nsys, int* out, const int* skip) {
    hipLaunchKernelGGL(fcd::k_count_running, dim3(1), dim3(1024), 0, st, status, nsys, out, skip);
    return hipGetLastError();
}

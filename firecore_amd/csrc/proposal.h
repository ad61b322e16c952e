// Proposal network: a bonded-graph message-passing MLP that predicts the
// gains of the FIRE force preconditioner (precond.h) from per-atom invariants
// of the force, the pi-orbital and the substrate force.
//
//   x_i (16)  = features(element, F_i, pi_i, G_i, bonded neighbours)
//   h1 = silu(W1 x + b1)                        16 -> 64
//   h2 = silu(W2 [h1_i ; sum_{j~i} h1_j] + b2)  128 -> 64
//   h3 = silu(W3 [h2_i ; sum_{j~i} h2_j] + b3)  128 -> 64
//   o  = W4 h3 + b4                             64 -> 2
//   star gain of node k        g_k = gstar  * sigmoid(o_k,0)
//   molecule gain of comp. c   g_c = grigid * sigmoid(mean_{i in c} o_i,1)
//
// Dense layers run on the matrix cores, BF16 weights and activations, FP32
// accumulation: the 128->64 layers with gfx950's v_mfma_f32_16x16x32_bf16
// (K = 32 per instruction), the 16->64 input layer with
// v_mfma_f32_16x16x16_bf16 (its K is 16); the 64->2 head and the
// activations run on the vector ALUs in FP32.  Activations live in LDS as
// bf16 rows of HS halfwords (64 + 4 padding, so that the 16 rows read by one
// MFMA operand load fall into distinct banks).
#pragma once
#include "kernels.h"
#include "precond.h"

namespace fcd {

constexpr int MLP_NIN = 16, MLP_NH = 64, MLP_NOUT = 2;
constexpr int MLP_HS = 68;  // LDS row stride of the activations (bf16 elements)

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8s __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even (inputs are finite)
__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }

__device__ __forceinline__ float silu(float z) { return z / (1.f + __expf(-z)); }
__device__ __forceinline__ float sigmoidf_(float z) { return 1.f / (1.f + __expf(-z)); }

// Per-atom input features (the network's 16 inputs), from the exact atom force
// F, the substrate (GridFF) part G of it, the atom's pi-vector p (zero for
// caps), the mean force Fb of its bonded neighbours, its element class, its
// degree, whether it is a node, and its height above the substrate origin.
// All are invariant under rotations about the substrate normal z.
__device__ __forceinline__ void mlp_features(V3 F, V3 G, V3 p, V3 Fb, int elem, int deg, bool node, float height,
                                             float* x) {
    const float nF = sqrtf(dot(F, F)), nG = sqrtf(dot(G, G)), nFb = sqrtf(dot(Fb, Fb));
    const V3 Fh = F * (1.f / (nF + 1e-12f));
    const V3 Gh = G * (1.f / (nG + 1e-12f));
    const V3 Fbh = Fb * (1.f / (nFb + 1e-12f));
#pragma unroll
    for (int k = 0; k < 4; k++) x[k] = (elem == k) ? 1.f : 0.f;
    const float fp = dot(Fh, p);
    x[4] = __logf(nF + 1e-4f);
    x[5] = Fh.z;
    x[6] = __logf(nG + 1e-4f);
    x[7] = Gh.z;
    x[8] = dot(Fh, Gh);
    x[9] = fp * fp;
    x[10] = p.z * p.z;
    x[11] = __logf(nFb + 1e-4f);
    x[12] = dot(Fh, Fbh);
    x[13] = 0.25f * (float)deg;
    x[14] = node ? 1.f : 0.f;
    x[15] = 0.1f * height;
}

// A operand of row block rb, K chunk kc from the activation rows in LDS.
// Chunks 0..KS-1 read the atom's own row; chunks KS.. read the sum over its
// bonded neighbours (FP32 sum of bf16 rows, rounded to bf16).
template <int KS>
__device__ __forceinline__ bf16x4 mlp_a_frag(const unsigned short* h, const int4* nb, int n, int rb, int kc,
                                             int lane) {
    const int row = rb * 16 + (lane & 15);
    const int kk = 4 * (lane >> 4);
    bf16x4 a;
    if (kc < KS) {
        // 4 consecutive bf16 = one 8-byte load (rows are 8-byte aligned); rows
        // past n repeat row n-1 (their results are discarded)
        const int rr = min(row, n - 1);
        const uint2 u = *reinterpret_cast<const uint2*>(h + rr * MLP_HS + kc * 16 + kk);
        a[0] = (short)(u.x & 0xFFFFu);
        a[1] = (short)(u.x >> 16);
        a[2] = (short)(u.y & 0xFFFFu);
        a[3] = (short)(u.y >> 16);
    } else {
        const int4 n4 = (row < n) ? nb[row] : make_int4(-1, -1, -1, -1);
        const int ns[4] = {n4.x, n4.y, n4.z, n4.w};
        float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            if (ns[t] < 0) continue;
            const uint2 u = *reinterpret_cast<const uint2*>(h + ns[t] * MLP_HS + (kc - KS) * 16 + kk);
            s0 += __uint_as_float(u.x << 16);
            s1 += __uint_as_float(u.x & 0xFFFF0000u);
            s2 += __uint_as_float(u.y << 16);
            s3 += __uint_as_float(u.y & 0xFFFF0000u);
        }
        a[0] = (short)f2bf(s0);
        a[1] = (short)f2bf(s1);
        a[2] = (short)f2bf(s2);
        a[3] = (short)f2bf(s3);
    }
    return a;
}

// The same for the 32-wide K chunks of v_mfma_f32_16x16x32_bf16 (lane l holds
// row l%16, K values 8*(l/16)..+7 of the chunk): chunks 0,1 are the atom's own
// row (64 columns), chunks 2,3 the bf16-rounded FP32 sum over its bonded
// neighbours, element by element as in mlp_a_frag.
__device__ __forceinline__ bf16x8 mlp_a_frag32(const unsigned short* h, const int4* nb, int n, int rb, int kc,
                                               int lane) {
    const int row = rb * 16 + (lane & 15);
    const int kk = 8 * (lane >> 4);
    bf16x8s a;
    if (kc < 2) {
        const int rr = min(row, n - 1);
        const unsigned short* q = h + rr * MLP_HS + kc * 32 + kk;  // rows are 8-byte aligned
        const uint2 u0 = *reinterpret_cast<const uint2*>(q);
        const uint2 u1 = *reinterpret_cast<const uint2*>(q + 4);
        a[0] = (short)(u0.x & 0xFFFFu);
        a[1] = (short)(u0.x >> 16);
        a[2] = (short)(u0.y & 0xFFFFu);
        a[3] = (short)(u0.y >> 16);
        a[4] = (short)(u1.x & 0xFFFFu);
        a[5] = (short)(u1.x >> 16);
        a[6] = (short)(u1.y & 0xFFFFu);
        a[7] = (short)(u1.y >> 16);
    } else {
        const int4 n4 = (row < n) ? nb[row] : make_int4(-1, -1, -1, -1);
        const int ns[4] = {n4.x, n4.y, n4.z, n4.w};
        float sm[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < 4; t++) {
            if (ns[t] < 0) continue;
            const unsigned short* q = h + ns[t] * MLP_HS + (kc - 2) * 32 + kk;
            const uint2 u0 = *reinterpret_cast<const uint2*>(q);
            const uint2 u1 = *reinterpret_cast<const uint2*>(q + 4);
            sm[0] += __uint_as_float(u0.x << 16);
            sm[1] += __uint_as_float(u0.x & 0xFFFF0000u);
            sm[2] += __uint_as_float(u0.y << 16);
            sm[3] += __uint_as_float(u0.y & 0xFFFF0000u);
            sm[4] += __uint_as_float(u1.x << 16);
            sm[5] += __uint_as_float(u1.x & 0xFFFF0000u);
            sm[6] += __uint_as_float(u1.y << 16);
            sm[7] += __uint_as_float(u1.y & 0xFFFF0000u);
        }
#pragma unroll
        for (int e = 0; e < 8; e++) a[e] = (short)f2bf(sm[e]);
    }
    return __builtin_bit_cast(bf16x8, a);
}

// B operand (W^T) of column block cb, 32-wide K chunk kc: W row-major [64][128].
__device__ __forceinline__ bf16x8 mlp_b_frag32(const unsigned short* W, int cb, int kc, int lane) {
    const unsigned short* q = W + (cb * 16 + (lane & 15)) * 128 + kc * 32 + 8 * (lane >> 4);
    const uint4 u = *reinterpret_cast<const uint4*>(q);  // 16-byte aligned: rows of 256 B
    bf16x8s b;
    b[0] = (short)(u.x & 0xFFFFu);
    b[1] = (short)(u.x >> 16);
    b[2] = (short)(u.y & 0xFFFFu);
    b[3] = (short)(u.y >> 16);
    b[4] = (short)(u.z & 0xFFFFu);
    b[5] = (short)(u.z >> 16);
    b[6] = (short)(u.w & 0xFFFFu);
    b[7] = (short)(u.w >> 16);
    return __builtin_bit_cast(bf16x8, b);
}

// B operand (W^T) of column block cb, K chunk kc: W row-major [64][KIN].
template <int KIN>
__device__ __forceinline__ bf16x4 mlp_b_frag(const unsigned short* W, int cb, int kc, int lane) {
    const unsigned short* q = W + (cb * 16 + (lane & 15)) * KIN + kc * 16 + 4 * (lane >> 4);
    const uint2 u = *reinterpret_cast<const uint2*>(q);
    bf16x4 b;
    b[0] = (short)(u.x & 0xFFFFu);
    b[1] = (short)(u.x >> 16);
    b[2] = (short)(u.y & 0xFFFFu);
    b[3] = (short)(u.y >> 16);
    return b;
}

// One dense layer over all rows (atoms) of a system: acc[q] = tile t = wid + q*NW
// (row block t/4, column block t%4) of [A] W^T, K = KIN (16: the 16 inputs in
// columns 0..15; 128: own row + neighbour sum).  At most 16 tiles per wave
// because rows <= BLOCK.
template <int BLOCK, int KIN>
__device__ __forceinline__ void mlp_layer(const unsigned short* h, const int4* nb, const unsigned short* W, int n,
                                         int nrb, f32x4* acc) {
    constexpr int NW = BLOCK / 64;
    constexpr int NK = KIN / 16;
    constexpr int KS = (KIN == 16) ? 1 : 4;
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
        const int t = wid + q * NW;
        if (t < nrb * 4) {
            const int rb = t >> 2, cb = t & 3;
            if constexpr (KIN == 128) {
#pragma unroll
                for (int kc = 0; kc < 4; kc++) {
                    const bf16x8 a = mlp_a_frag32(h, nb, n, rb, kc, lane);
                    const bf16x8 b = mlp_b_frag32(W, cb, kc, lane);
                    acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[q], 0, 0, 0);
                }
            } else {
#pragma unroll
                for (int kc = 0; kc < NK; kc++) {
                    const bf16x4 a = mlp_a_frag<KS>(h, nb, n, rb, kc, lane);
                    const bf16x4 b = mlp_b_frag<KIN>(W, cb, kc, lane);
                    acc[q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[q], 0, 0, 0);
                }
            }
        }
    }
}

// Bias + SiLU of the accumulated tiles, written back into the activation rows
// (rows >= n are left as they are; they are never read as neighbours).
template <int BLOCK>
__device__ __forceinline__ void mlp_store(unsigned short* h, const float* bias, int nrb, int n, const f32x4* acc,
                                         float* zout, int zstride) {
    constexpr int NW = BLOCK / 64;
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int t = wid + q * NW;
        if (t < nrb * 4) {
            const int rb = t >> 2, cb = t & 3;
            const int col = cb * 16 + (lane & 15);
            const float bc = bias[col];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int row = rb * 16 + 4 * (lane >> 4) + i;
                const float z = acc[q][i] + bc;
                if (row < n) {
                    h[row * MLP_HS + col] = f2bf(silu(z));
                    if (zout) zout[row * zstride + col] = z;
                }
            }
        }
    }
}

// Forward pass for one system whose features (bf16) are in columns 0..15 of
// the activation rows h[0..n).  nb = bonded neighbours (system-local).  On
// return out[i*2 + j] = o_ij for i < n.  z (optional, [3][n][64] FP32) and hs
// (optional, [3][n][64] bf16 rows h1, h2, h3) receive the pre-activations and
// activations for the backward pass.  Every thread of the block calls it.
template <int BLOCK>
__device__ __forceinline__ void mlp_forward(const MlpDev& M, unsigned short* h, const int4* nb, int n, float* out, float* z,
                            unsigned short* hs) {
    const int nrb = (n + 15) >> 4;
    f32x4 acc[16];
    const int zs = MLP_NH;
    mlp_layer<BLOCK, 16>(h, nb, M.W1, n, nrb, acc);
    __syncthreads();
    mlp_store<BLOCK>(h, M.b1, nrb, n, acc, z, zs);
    __syncthreads();
    if (hs)
        for (int e = threadIdx.x; e < n * MLP_NH; e += BLOCK) hs[e] = h[(e >> 6) * MLP_HS + (e & 63)];
    mlp_layer<BLOCK, 128>(h, nb, M.W2, n, nrb, acc);
    __syncthreads();
    mlp_store<BLOCK>(h, M.b2, nrb, n, acc, z ? z + n * MLP_NH : nullptr, zs);
    __syncthreads();
    if (hs)
        for (int e = threadIdx.x; e < n * MLP_NH; e += BLOCK) hs[n * MLP_NH + e] = h[(e >> 6) * MLP_HS + (e & 63)];
    mlp_layer<BLOCK, 128>(h, nb, M.W3, n, nrb, acc);
    __syncthreads();
    mlp_store<BLOCK>(h, M.b3, nrb, n, acc, z ? z + 2 * n * MLP_NH : nullptr, zs);
    __syncthreads();
    if (hs)
        for (int e = threadIdx.x; e < n * MLP_NH; e += BLOCK) hs[2 * n * MLP_NH + e] = h[(e >> 6) * MLP_HS + (e & 63)];
    for (int i = threadIdx.x; i < n; i += BLOCK) {
        float o0 = M.b4[0], o1 = M.b4[1];
        const unsigned short* q = h + i * MLP_HS;
#pragma unroll 8
        for (int k = 0; k < MLP_NH; k++) {
            const float v = bf2f(q[k]);
            o0 += M.W4[k] * v;
            o1 += M.W4[MLP_NH + k] * v;
        }
        out[i * 2] = o0;
        out[i * 2 + 1] = o1;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Forward pass used by the relaxation kernel between segments of its step
// loop (nothing of the force loop is live there, so a layer may hold all its
// weight fragments in registers).  Activations alternate between two
// row buffers of n rows, hA (LDS) and hG (LDS when the workgroup's LDS budget
// allows, else a per-workgroup slot in global memory that stays in L2):
// layer 1 hA -> hG, layer 2 hG -> hA, layer 3 hA -> hG, head from hG.
template <int BLOCK, int KIN>
__device__ __forceinline__ void mlp_layer_rows(const unsigned short* hin, const int4* nb, const unsigned short* W,
                                               const float* bias, int n, unsigned short* hout) {
    constexpr int NW = BLOCK / 64;
    constexpr int NK = KIN / 16;
    constexpr int KS = (KIN == 16) ? 1 : 4;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nrb = (n + 15) >> 4;
    // the weight fragments of all four column blocks are loaded once per layer
    // (one memory latency); a wave then takes whole row blocks, builds their A
    // operand (own rows and neighbour sums) once and runs four independent
    // accumulation chains
    // K = 32 chunks for the 128-wide layers, K = 16 for the input layer
    constexpr int NK32 = (KIN == 128) ? 4 : 1;
    constexpr int NK16 = (KIN == 128) ? 1 : NK;
    bf16x8 bw8[4][NK32];
    bf16x4 bw[4][NK16];
#pragma unroll
    for (int cb = 0; cb < 4; cb++) {
        if constexpr (KIN == 128) {
#pragma unroll
            for (int kc = 0; kc < NK32; kc++) bw8[cb][kc] = mlp_b_frag32(W, cb, kc, lane);
        } else {
#pragma unroll
            for (int kc = 0; kc < NK16; kc++) bw[cb][kc] = mlp_b_frag<KIN>(W, cb, kc, lane);
        }
    }
    float bc[4];
#pragma unroll
    for (int cb = 0; cb < 4; cb++) bc[cb] = bias[cb * 16 + (lane & 15)];
#pragma unroll 1
    for (int rb = wid; rb < nrb; rb += NW) {
        f32x4 acc[4];
#pragma unroll
        for (int cb = 0; cb < 4; cb++) acc[cb] = f32x4{0.f, 0.f, 0.f, 0.f};
        if constexpr (KIN == 128) {
            bf16x8 a[NK32];
#pragma unroll
            for (int kc = 0; kc < NK32; kc++) a[kc] = mlp_a_frag32(hin, nb, n, rb, kc, lane);
#pragma unroll
            for (int kc = 0; kc < NK32; kc++)
#pragma unroll
                for (int cb = 0; cb < 4; cb++)
                    acc[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[kc], bw8[cb][kc], acc[cb], 0, 0, 0);
        } else {
            bf16x4 a[NK16];
#pragma unroll
            for (int kc = 0; kc < NK16; kc++) a[kc] = mlp_a_frag<KS>(hin, nb, n, rb, kc, lane);
#pragma unroll
            for (int kc = 0; kc < NK16; kc++)
#pragma unroll
                for (int cb = 0; cb < 4; cb++)
                    acc[cb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[kc], bw[cb][kc], acc[cb], 0, 0, 0);
        }
#pragma unroll
        for (int cb = 0; cb < 4; cb++) {
            const int col = cb * 16 + (lane & 15);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int row = rb * 16 + 4 * (lane >> 4) + i;
                if (row < n) hout[row * MLP_HS + col] = f2bf(silu(acc[cb][i] + bc[cb]));
            }
        }
    }
}

// Forward pass with the features in columns 0..15 of hA (n <= BLOCK rows);
// returns the two head outputs of row threadIdx.x (zeros for threads >= n).
// The global writes of a layer are made visible to the whole workgroup by
// the fence + barrier that follow them.
template <int BLOCK>
__device__ __forceinline__ float2 mlp_forward_rows(const MlpDev& M, unsigned short* hA, unsigned short* hG,
                                                   const int4* nb, int n) {
    mlp_layer_rows<BLOCK, 16>(hA, nb, M.W1, M.b1, n, hG);
    __threadfence_block();
    __syncthreads();
    mlp_layer_rows<BLOCK, 128>(hG, nb, M.W2, M.b2, n, hA);
    __syncthreads();
    mlp_layer_rows<BLOCK, 128>(hA, nb, M.W3, M.b3, n, hG);
    __threadfence_block();
    __syncthreads();
    float2 o = make_float2(0.f, 0.f);
    const int i = threadIdx.x;
    if (i < n) {
        o = make_float2(M.b4[0], M.b4[1]);
        const unsigned short* q = hG + i * MLP_HS;
#pragma unroll 8
        for (int k = 0; k < MLP_NH; k++) {
            const float v = bf2f(q[k]);
            o.x += M.W4[k] * v;
            o.y += M.W4[MLP_NH + k] * v;
        }
    }
    __syncthreads();
    return o;
}

// Feature row of atom i (bf16, columns 0..15 of its activation row).  F is
// the exact atom force, G its substrate part (recomputed here), pi the atom's
// pi-vector (nodes only), Fb the mean exact force of its bonded neighbours.
__device__ __forceinline__ void mlp_feature_row(const RunParams& P, const float4* s_pos, const float4* s_F,
                                                const int4* nb, const int* elem, int N, int M, int i, float4 plq,
                                                unsigned short* hrow, const V3* Gpre = nullptr) {
    const V3 p = xyz(s_pos[i]);
    V3 G = zero3();
    if (Gpre) G = *Gpre;  // substrate force already sampled at this geometry
    else if (P.bGridFF) grid_sample(P.grid, p, mk(plq.x, plq.y, plq.z), G);
    const int4 n4 = nb[i];
    const int ns[4] = {n4.x, n4.y, n4.z, n4.w};
    V3 Fb = zero3();
    int deg = 0;
#pragma unroll
    for (int t = 0; t < 4; t++)
        if (ns[t] >= 0) {
            Fb += xyz(s_F[ns[t]]);
            deg++;
        }
    if (deg) Fb = Fb * (1.f / (float)deg);
    const V3 pi = (i < M) ? xyz(s_pos[N + i]) : zero3();
    const float height = P.bGridFF ? p.z - P.grid.p0z : 0.f;
    float x[MLP_NIN];
    mlp_features(xyz(s_F[i]), G, pi, Fb, elem[i], deg, i < M, height, x);
#pragma unroll
    for (int k = 0; k < MLP_NIN; k++) hrow[k] = f2bf(x[k]);
}

// Network evaluation inside the relaxation: features from the exact forces in
// s_F (the substrate force of each atom is sampled again here), forward pass,
// gains into s_gain (per node, then per component at s_gain[M + c]).  s_cred
// receives per-wave partial sums (component output sums and atom counts).
// Every thread calls it; nothing of the step loop is live across the call.
template <int BLOCK>
__device__ __forceinline__ void propose_gains(const RunParams& P, const int4* nb, const float4* req, const int* elem,
                              const float4* s_pos, const float4* s_F, float4* s_req, unsigned short* h,
                              unsigned short* hG, float* s_gain, float* s_cred, int N, int M, int nc, int myc,
                              float4 plq) {
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wid = tid >> 6;
    constexpr int NW = BLOCK / 64;
    // the bonded neighbour lists are read for every operand fragment: stage
    // them in LDS over the (constant) REQ rows, restored from req below
    int4* s_nb = reinterpret_cast<int4*>(s_req);
    if (tid < N) {
        mlp_feature_row(P, s_pos, s_F, nb, elem, N, M, tid, plq, h + tid * MLP_HS);
        s_nb[tid] = nb[tid];
    }
    __syncthreads();
    const float2 o = mlp_forward_rows<BLOCK>(P.mlp, h, hG, s_nb, N);
    if (tid < N) s_req[tid] = req[tid];
    if (tid < M) s_gain[tid] = P.pc_gstar * sigmoidf_(o.x);
    for (int c = 0; c < nc; c++) {
        const bool mine = tid < N && myc == c;
        const float v = wave_sum(mine ? o.y : 0.f);
        const float n = wave_sum(mine ? 1.f : 0.f);
        if (lane == 0) {
            s_cred[(wid * PC_MAXC + c) * PC_NSUM] = v;
            s_cred[(wid * PC_MAXC + c) * PC_NSUM + 1] = n;
        }
    }
    __syncthreads();
    if (tid < nc) {
        float s = 0.f, n = 0.f;
#pragma unroll
        for (int w = 0; w < NW; w++) {
            s += s_cred[(w * PC_MAXC + tid) * PC_NSUM];
            n += s_cred[(w * PC_MAXC + tid) * PC_NSUM + 1];
        }
        s_gain[M + tid] = P.pc_grigid * sigmoidf_(s / fmaxf(n, 1.f));
    }
    __syncthreads();
}

}  // namespace fcd

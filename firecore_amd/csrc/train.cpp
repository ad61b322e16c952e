// Host side of the proposal network: parameter installation, trajectory
// snapshots, and the device training step (kernels in proposal_train.hip).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "precond.h"
#include "world.h"

namespace fcd {
struct TrainBufs {
    const float4* x0;
    const float4* F0;
    const double* E0;
    unsigned short* X;
    float* z;
    unsigned short* hs;
    float* out;
    float* dout;
    float* gnode;
    float* gcomp;
    const float4* pcforce;
    float* loss;
    float* gpart;
    float eta, dmax;
};
}  // namespace fcd

extern "C" int fc_mlp_nparam();
extern "C" hipError_t fc_launch_train_fwd(hipStream_t st, const fcd::BatchDev* B, const fcd::RunParams* P,
                                          const fcd::TrainBufs* T);
extern "C" hipError_t fc_launch_train_displace(hipStream_t st, const fcd::BatchDev* B, const fcd::TrainBufs* T);
extern "C" hipError_t fc_launch_train_loss_grad(hipStream_t st, const fcd::BatchDev* B, const fcd::RunParams* P,
                                                const fcd::TrainBufs* T);
extern "C" hipError_t fc_launch_train_bwd(hipStream_t st, const fcd::BatchDev* B, const fcd::RunParams* P,
                                          const fcd::TrainBufs* T);
extern "C" hipError_t fc_launch_grad_reduce(hipStream_t st, const float* gpart, int nsys, float* grad,
                                            const float* loss, float* loss_out);
extern "C" hipError_t fc_launch_adam(hipStream_t st, float* theta, float* m, float* v, const float* grad, float lr,
                                     float b1, float b2, float eps, float bc1, float bc2, float clip,
                                     unsigned short* Wbf, float* Ff);
extern "C" hipError_t fc_launch_precond(hipStream_t st, const fcd::BatchDev* B, const fcd::RunParams* P, float4* out,
                                        const float* gnode, const float* gcomp);
extern "C" hipError_t fc_launch_gather_snap(hipStream_t st, const fcd::BatchDev* B, const float4* snap,
                                            const int* idx);
extern "C" hipError_t fc_launch_mlp_gains(hipStream_t st, const fcd::BatchDev* B, const fcd::RunParams* P, float* gnode,
                                          float* gcomp);
extern "C" hipError_t fc_launch_zero_int(hipStream_t st, int* a, int n);
extern "C" hipError_t fc_launch_count_running(hipStream_t st, const int* status, int nsys, int* out,
                                              const int* skip);
extern "C" hipError_t fc_launch_zero_f4(hipStream_t st, float4* a, int n);
extern "C" hipError_t fc_launch_relax_persistent(int block, int hb, int ngroups, size_t lds_bytes, hipStream_t stream,
                                                 const fcd::BatchDev* B, const fcd::RunParams* P, const int* order,
                                                 const int* qinfo, int* counters, int cls, int nmax, int mmax,
                                                 int wmax, int steal, unsigned short* hslot, int h2lds);

namespace fcw {

static fcd::TrainBufs train_bufs(World& W, float eta, float dmax) {
    fcd::TrainBufs T{};
    T.x0 = W.d_x0.p;
    T.F0 = W.d_F0.p;
    T.E0 = W.d_E0.p;
    T.X = W.d_X.p;
    T.z = W.d_z.p;
    T.hs = W.d_hs.p;
    T.out = W.d_out.p;
    T.dout = W.d_dout.p;
    T.gnode = W.d_gnode.p;
    T.gcomp = W.d_gcomp.p;
    T.pcforce = W.d_pcforce.p;
    T.loss = W.d_loss.p;
    T.gpart = W.d_gpart.p;
    T.eta = eta;
    T.dmax = dmax;
    return T;
}

bool World::record_snapshots(int every, int n) {
    if (!dev_ready || nSys == 0 || every < 1 || n < 0) return false;
    snap_every = every;
    nsnap = n;
    snap_on = n > 0;
    if (snap_on) d_snap.alloc((size_t)n * ntotV);
    return true;
}

bool World::load_snapshots(const std::vector<int>& idx) {
    if (!snap_on || (int)idx.size() != nSys) return false;
    for (int k : idx)
        if (k < 0 || k >= nsnap) return false;
    d_snapidx.upload(idx, stream);
    fcd::BatchDev B = batch_dev();
    return hip_ok(fc_launch_gather_snap(stream, &B, d_snap.p, d_snapidx.p), "gather snapshots");
}

bool World::prop_alloc() {
    if (!dev_ready || nSys == 0) return false;
    if (maxN > 256 || maxM > 256) {
        std::fprintf(stderr, "[libMMFFmulti_lib] proposal network: systems above 256 atoms are not supported\n");
        return false;
    }
    if (!d_mlpW.p) {
        std::fprintf(stderr, "[libMMFFmulti_lib] proposal network: no parameters installed\n");
        return false;
    }
    const int NP = fc_mlp_nparam();
    d_x0.alloc(ntotV);
    d_F0.alloc(ntotV);
    d_pcforce.alloc(ntotV);
    d_E0.alloc(nSys);
    d_X.alloc((size_t)ntotA * 16);
    d_z.alloc((size_t)ntotA * 3 * 64);
    d_hs.alloc((size_t)ntotA * 3 * 64);
    d_out.alloc((size_t)ntotA * 2);
    d_dout.alloc((size_t)ntotA * 2);
    d_gnode.alloc(std::max(ntotM, 1));
    d_gcomp.alloc((size_t)nSys * fcd::PC_MAXC);
    d_loss.alloc(nSys);
    d_gpart.alloc((size_t)nSys * NP);
    d_grad.alloc(NP);
    d_lossmean.alloc(1);
    if (!theta_dev) {
        d_theta.upload(mlp_theta, stream);
        std::vector<float> z(NP, 0.f);
        d_am.upload(z, stream);
        d_av.upload(z, stream);
        adam_t = 0;
        theta_dev = true;
    }
    return true;
}

// exact forces and energies at the current positions become the reference
// state (x0, F0, E0) of the training step
bool World::prop_prepare() {
    if (!prop_alloc() || !eval_forces()) return false;
    hip_ok(hipMemcpyAsync(d_x0.p, d_atoms.p, sizeof(float4) * ntotV, hipMemcpyDeviceToDevice, stream), "x0");
    hip_ok(hipMemcpyAsync(d_F0.p, d_aforce.p, sizeof(float4) * ntotV, hipMemcpyDeviceToDevice, stream), "F0");
    return hip_ok(hipMemcpyAsync(d_E0.p, d_Esys.p, sizeof(double) * nSys, hipMemcpyDeviceToDevice, stream), "E0");
}

bool World::prop_forward() {
    fcd::BatchDev B = batch_dev();
    fcd::RunParams P = run_params(0.0, 0);
    fcd::TrainBufs T = train_bufs(*this, 0.f, 0.f);
    return hip_ok(fc_launch_train_fwd(stream, &B, &P, &T), "proposal forward");
}

bool World::prop_loss_grad(float eta, float dmax) {
    fcd::BatchDev B = batch_dev();
    fcd::RunParams P = run_params(0.0, 0);
    fcd::TrainBufs T = train_bufs(*this, eta, dmax);
    if (!hip_ok(fc_launch_precond(stream, &B, &P, d_pcforce.p, d_gnode.p, d_gcomp.p), "precond (learned)"))
        return false;
    if (!hip_ok(fc_launch_train_displace(stream, &B, &T), "displace")) return false;
    if (!eval_forces()) return false;
    if (!hip_ok(fc_launch_train_loss_grad(stream, &B, &P, &T), "loss gradient")) return false;
    // the batch returns to the snapshot positions
    return hip_ok(hipMemcpyAsync(d_atoms.p, d_x0.p, sizeof(float4) * ntotV, hipMemcpyDeviceToDevice, stream),
                  "restore x0");
}

bool World::prop_backward() {
    fcd::BatchDev B = batch_dev();
    fcd::RunParams P = run_params(0.0, 0);
    fcd::TrainBufs T = train_bufs(*this, 0.f, 0.f);
    if (!hip_ok(fc_launch_train_bwd(stream, &B, &P, &T), "proposal backward")) return false;
    return hip_ok(fc_launch_grad_reduce(stream, d_gpart.p, nSys, d_grad.p, d_loss.p, d_lossmean.p), "grad reduce");
}

bool World::prop_adam(float lr) {
    adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.f - std::pow(b1, (float)adam_t), bc2 = 1.f - std::pow(b2, (float)adam_t);
    return hip_ok(fc_launch_adam(stream, d_theta.p, d_am.p, d_av.p, d_grad.p, lr, b1, b2, 1e-8f, bc1, bc2, 1e3f,
                                 d_mlpW.p, d_mlpF.p),
                  "adam");
}

// Relaxation with learned preconditioner gains: the persistent kernel runs
// in chunks of pc_every steps per system (state suspended to and resumed
// from global memory); between chunks the proposal network re-evaluates the
// gains of every running system.  One chunk iteration (network, queue reset,
// the size-class launches) is captured once as a graph and replayed; every
// `check` iterations the number of running systems is read back.
int World::relax_learned(int nstep_max, double Fconv, bool restore_start) {
    if (!dev_ready || nSys == 0) return -1;
    if (!d_mlpW.p) {
        std::fprintf(stderr, "[libMMFFmulti_lib] learned relaxation: no proposal weights installed\n");
        return -1;
    }
    if (!ring_ok) {
        std::fprintf(stderr, "[libMMFFmulti_lib] learned relaxation needs symmetric exclusions\n");
        return -1;
    }
    d_gnode.alloc(std::max(ntotM, 1));
    d_gcomp.alloc((size_t)nSys * fcd::PC_MAXC);
    d_count.alloc(1);
    if (autotune) plan_classes();
    if (restore_start)
        hip_ok(hipMemcpyAsync(d_atoms.p, d_atoms0.p, sizeof(float4) * ntotV, hipMemcpyDeviceToDevice, stream),
               "restore");
    // systems too large for every class are relaxed afterwards (relax_unclassed)
    // and not counted by the chunk loop
    if (n_unclassed > 0) {
        std::vector<int> skip(nSys);
        for (int s = 0; s < nSys; s++) skip[s] = sys_class[s] < 0 ? 1 : 0;
        d_only.upload(skip, stream);
    }
    EventSet events;
    hipEvent_t e0 = events.make(), e1 = events.make();
    hipEventRecord(e0, stream);
    // exact forces at the start (first network evaluation); resets status, steps and FIRE state
    if (!eval_forces()) return -1;
    fc_launch_zero_f4(stream, d_avel.p, ntotV);
    fc_launch_zero_f4(stream, d_stall.p, nSys);  // fallback flags of this relaxation only
    fcd::BatchDev B = batch_dev();
    fcd::RunParams P = run_params(Fconv, nstep_max);
    P.chunk = std::max(pc_every, 1);
    // events for the fork/join of the class launches
    std::vector<hipEvent_t> evs(classes.size() + 1);
    for (auto& e : evs) e = events.make(hipEventDisableTiming);
    auto iteration = [&](hipStream_t st) -> bool {
        if (!hip_ok(fc_launch_mlp_gains(st, &B, &P, d_gnode.p, d_gcomp.p), "gains")) return false;
        if (!hip_ok(fc_launch_zero_int(st, d_queue.p, 8), "queue reset")) return false;
        hipEventRecord(evs[0], st);
        for (size_t k = 0; k < classes.size(); k++) {
            SizeClass& C = classes[k];
            hipStream_t cs = class_streams[k % class_streams.size()];
            hipStreamWaitEvent(cs, evs[0], 0);
            if (!hip_ok(fc_launch_relax_persistent(C.block, hb_batch ? 1 : 0, C.ngroups, C.lds, cs, &B, &P, d_order.p,
                                                   qinfo.data(), d_queue.p, C.cls, C.nmax, C.mmax, C.wmax,
                                                   sched_mode == 0 ? 1 : 0, nullptr, 0),
                        "relax chunk"))
                return false;
            hipEventRecord(evs[k + 1], cs);
            hipStreamWaitEvent(st, evs[k + 1], 0);
        }
        return true;
    };
    const int check = 16;
    struct GraphGuard {  // the captured chunk, released on every exit path
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        ~GraphGuard() {
            if (exec) hipGraphExecDestroy(exec);
            if (graph) hipGraphDestroy(graph);
        }
    } cg;
    bool use_graph = use_graphs;
    if (use_graph) {
        use_graph = hipStreamBeginCapture(stream, hipStreamCaptureModeGlobal) == hipSuccess;
        if (use_graph) {
            const bool ok = iteration(stream);
            use_graph = (hipStreamEndCapture(stream, &cg.graph) == hipSuccess) && ok && cg.graph;
            if (use_graph) use_graph = hipGraphInstantiate(&cg.exec, cg.graph, nullptr, nullptr, 0) == hipSuccess;
        }
        if (!use_graph) {
            (void)hipGetLastError();
            std::fprintf(stderr, "[libMMFFmulti_lib] graph capture unavailable; launching the chunks directly\n");
        }
    }
    std::vector<int> cnt(1, 1);
    const int max_iter = nstep_max / P.chunk + 2;
    int it = 0;
    bool ok = true;
    for (; it < max_iter && ok;) {
        for (int k = 0; k < check && it < max_iter; k++, it++) {
            if (use_graph) ok = hip_ok(hipGraphLaunch(cg.exec, stream), "graph launch");
            else ok = iteration(stream);
            if (!ok) break;
        }
        fc_launch_count_running(stream, d_status.p, nSys, d_count.p, n_unclassed > 0 ? d_only.p : nullptr);
        d_count.download(cnt, stream);
        if (!hip_ok(hipStreamSynchronize(stream), "learned relax")) ok = false;
        if (cnt[0] == 0) break;
    }
    hipEventRecord(e1, stream);
    hipStreamSynchronize(stream);
    float ms = 0;
    hipEventElapsedTime(&ms, e0, e1);
    last_ms = ms;
    last_chunks = it;
    if (!ok) return -1;
    if (n_unclassed > 0 && !relax_unclassed(nstep_max, Fconv)) return -1;
    d_nsteps.download(nsteps, stream);
    d_status.download(status, stream);
    if (!hip_ok(hipStreamSynchronize(stream), "learned relax results")) return -1;
    int mx = 0;
    for (int s = 0; s < nSys; s++) mx = std::max(mx, nsteps[s]);
    return mx;
}

bool World::restore_positions() {
    if (!dev_ready || !d_atoms0.p) return false;
    hip_ok(hipMemcpyAsync(d_atoms.p, d_atoms0.p, sizeof(float4) * ntotV, hipMemcpyDeviceToDevice, stream), "restore");
    return hip_ok(hipStreamSynchronize(stream), "restore");
}

bool World::download_theta() {
    if (!theta_dev) return true;
    mlp_theta.resize(fc_mlp_nparam());
    d_theta.download(mlp_theta, stream);
    return hip_ok(hipStreamSynchronize(stream), "download theta");
}

}  // namespace fcw

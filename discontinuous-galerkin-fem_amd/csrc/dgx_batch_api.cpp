// Batched instances: per-instance contexts, the work-item list, the graph-captured two-step
// window of persistent stage launches, sequential comparison and the tile autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "dgx_ctx.hpp"

struct dgx_batch {
    std::vector<dgx_ctx*> inst;
    int precision = DGX_FP32;
    int order_mask = 0;
    int cus = 256;
    hipStream_t stream = nullptr;
    dgx_batch_tuning tune{2, {4, 4, 4, 4}, 4};
    DevBuf params[6];                 // [window stage][instance] StageParams
    DevBuf items_all, items_p[4];     // work lists (fused, per order)
    int n_all = 0, n_p[4] = {0, 0, 0, 0};
    const void* tabs[4] = {nullptr, nullptr, nullptr, nullptr};
    // orders whose instances all run the matrix-core stage (FP32, p = 3, 4): their lane tables;
    // those items always go through the per-order matrix-core launch
    const void* tabm[4] = {nullptr, nullptr, nullptr, nullptr};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // layout 2: one stream per order (default priority), forked from and joined back into
    // `stream` around each run of window stages, so that one order's launch tail overlaps the
    // next order's launches
    hipStream_t ostream[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t fork = nullptr, join[4] = {nullptr, nullptr, nullptr, nullptr};
    bool dirty = true;

    ~dgx_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (fork) (void)hipEventDestroy(fork);
        for (int p = 0; p < 4; ++p) {
            if (join[p]) (void)hipEventDestroy(join[p]);
            if (ostream[p]) (void)hipStreamDestroy(ostream[p]);
        }
        for (auto* c : inst) dgx_destroy(c);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Stage parameters of one step of the window.  Step 0 reads u and ends in u1, step 1 reads
    // u1 and ends in u, so replaying the two-step window leaves every state in place.
    StageParams stage_params(dgx_ctx* c, int step, int stage) {
        void* a = step == 0 ? c->u.p : c->u1.p;  // u^n of this step
        void* b = step == 0 ? c->u1.p : c->u.p;  // result of this step
        void* t = c->u2.p;
        StageParams sp{};
        const void* in[3] = {a, b, t};
        void* outb[3] = {b, t, b};
        const double ca[3] = {0.0, 0.75, 1.0 / 3.0}, cb[3] = {1.0, 0.25, 2.0 / 3.0};
        sp.uk = in[stage];
        sp.un = a;
        sp.out = outb[stage];
        sp.geo = c->geo.as<double>();
        sp.nbr = c->nbr.as<int32_t>();
        sp.code = c->code.as<uint32_t>();
        sp.tab = c->tabL.p;
        sp.bc = c->bc.as<double>();
        sp.K = c->K;
        sp.Kp = c->Kp;
        sp.dt = c->dt;
        sp.ca = ca[stage];
        sp.cb = cb[stage];
        sp.gamma = c->gamma;
        sp.mu = c->mu;
        sp.prandtl = c->prandtl;
        sp.sigma0 = c->sigma0;
        sp.mode = MODE_RK;
        sp.ns_face_quirk = 0;
        for (int v = 0; v < 4; ++v) sp.uref[v] = c->uref.v[v];
        sp.tabm = c->tabML.p;  // the matrix-core tile kernel's lane tables (null without them)
        c->set_tile_params(sp, MODE_RK);
        return sp;
    }

    void build() {
        const int n = (int)inst.size();
        for (int w = 0; w < 6; ++w) {
            std::vector<StageParams> h(n);
            for (int i = 0; i < n; ++i) h[i] = stage_params(inst[i], w / 3, w % 3);
            if (!params[w].p) params[w].alloc(n * sizeof(StageParams));
            hip_check(hipMemcpy(params[w].p, h.data(), n * sizeof(StageParams), hipMemcpyHostToDevice), "params");
        }
        // work items, most expensive orders first so the fixed-stride walk spreads them; items
        // of matrix-core orders are whole 64-element tiles and stay out of the fused list
        std::vector<BatchItem> all, per[4];
        for (int p = 4; p >= 1; --p)
            for (int i = 0; i < n; ++i) {
                dgx_ctx* c = inst[i];
                if (c->p != p) continue;
                const int chunk = 64 * std::max(1, tune.chunk[p - 1]);
                for (int e0 = 0; e0 < c->K; e0 += chunk) {
                    BatchItem it{i, e0, std::min(chunk, c->K - e0), p};
                    if (!tabm[p - 1]) all.push_back(it);
                    per[p - 1].push_back(it);
                }
            }
        auto upload = [](DevBuf& b, const std::vector<BatchItem>& v) {
            b.release();
            if (v.empty()) return;
            b.alloc(v.size() * sizeof(BatchItem));
            hip_check(hipMemcpy(b.p, v.data(), v.size() * sizeof(BatchItem), hipMemcpyHostToDevice), "items");
        };
        upload(items_all, all);
        n_all = (int)all.size();
        for (int p = 0; p < 4; ++p) {
            upload(items_p[p], per[p]);
            n_p[p] = (int)per[p].size();
        }
        // capture the two-step window
        if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
        if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
        hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture");
        launch_window(0, 6);
        hip_check(hipStreamEndCapture(stream, &graph), "end capture");
        hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "instantiate");
        dirty = false;
    }

    int grid() const { return std::max(1, cus * std::max(1, tune.blocks_per_cu)); }
    // the largest external-face trace count of a tile among the order-p instances
    int halo_cap(int p) const {
        int cap = 0;
        for (auto* c : inst)
            if (c->p == p) cap = std::max(cap, c->halo_cap);
        return cap;
    }

    // the order-p launch of window stage w on stream s
    void launch_order_stage(int w, int p, hipStream_t s) {
        const StageParams* pr = static_cast<const StageParams*>(params[w].p);
        if (tabm[p - 1])
            hip_check(launch_batch_stage_mfma(p, pr, items_p[p - 1].as<BatchItem>(), n_p[p - 1], halo_cap(p), grid(), s),
                      "batch stage (matrix cores)");
        else
            hip_check(launch_batch_stage_order(precision, p, pr, items_p[p - 1].as<BatchItem>(), n_p[p - 1], tabs[p - 1],
                                               grid(), s),
                      "batch stage (order)");
    }

    // window stages [w0, w1) on `stream` (captured or not); layout 2 forks one stream per order
    void launch_window(int w0, int w1) {
        if (tune.layout != 2) {
            for (int w = w0; w < w1; ++w) launch_window_stage(w);
            return;
        }
        hip_check(hipEventRecord(fork, stream), "fork");
        for (int p = 4; p >= 1; --p) {
            if (!n_p[p - 1]) continue;
            hip_check(hipStreamWaitEvent(ostream[p - 1], fork, 0), "fork wait");
            for (int w = w0; w < w1; ++w) launch_order_stage(w, p, ostream[p - 1]);
            hip_check(hipEventRecord(join[p - 1], ostream[p - 1]), "join");
            hip_check(hipStreamWaitEvent(stream, join[p - 1], 0), "join wait");
        }
    }

    void launch_window_stage(int w) {
        const StageParams* pr = static_cast<const StageParams*>(params[w].p);
        for (int p = 4; p >= 1; --p)
            if (tabm[p - 1] && n_p[p - 1]) launch_order_stage(w, p, stream);
        if (tune.layout == 0) {
            int mask = order_mask;
            for (int p = 1; p <= 4; ++p)
                if (tabm[p - 1]) mask &= ~(1 << (p - 1));
            if (n_all)
                hip_check(launch_batch_stage(precision, pr, items_all.as<BatchItem>(), n_all, tabs, mask, grid(), stream),
                          "batch stage");
        } else {
            for (int p = 4; p >= 1; --p)
                if (!tabm[p - 1] && n_p[p - 1]) launch_order_stage(w, p, stream);
        }
    }

    // nsteps steps; an odd remainder runs the first half of the window directly and swaps the
    // state buffers so that the window's pointers stay valid.
    void steps(int nsteps) {
        if (dirty) build();
        for (int s = 0; s + 1 < nsteps; s += 2) hip_check(hipGraphLaunch(exec, stream), "graph launch");
        if (nsteps & 1) {
            launch_window(0, 3);
            hip_check(hipStreamSynchronize(stream), "sync");
            for (auto* c : inst) std::swap(c->u, c->u1);
            dirty = true;
        }
    }

    std::vector<DevBuf> save() {
        std::vector<DevBuf> s(inst.size());
        for (size_t i = 0; i < inst.size(); ++i) {
            s[i].alloc(inst[i]->state_bytes());
            hip_check(hipMemcpyAsync(s[i].p, inst[i]->u.p, inst[i]->state_bytes(), hipMemcpyDeviceToDevice, stream), "save");
        }
        hip_check(hipStreamSynchronize(stream), "sync");
        return s;
    }
    void restore(const std::vector<DevBuf>& s) {
        for (size_t i = 0; i < inst.size(); ++i)
            hip_check(hipMemcpyAsync(inst[i]->u.p, s[i].p, inst[i]->state_bytes(), hipMemcpyDeviceToDevice, stream), "restore");
        hip_check(hipStreamSynchronize(stream), "sync");
    }

    double time_batched(int nsteps) {
        nsteps += nsteps & 1;
        if (dirty) build();
        steps(10);  // warm-up (two steps left the first measurement of a process some 3% slow)
        hip_check(hipStreamSynchronize(stream), "sync");
        hip_check(hipEventRecord(ev0, stream), "record");
        steps(nsteps);
        hip_check(hipEventRecord(ev1, stream), "record");
        hip_check(hipEventSynchronize(ev1), "sync");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, ev0, ev1), "elapsed");
        return ms / nsteps;
    }

    double time_sequential(int nsteps) {
        auto t0 = std::chrono::steady_clock::now();
        for (auto* c : inst) {
            for (int s = 0; s < nsteps; ++s) c->rk3_step(c->dt);
            hip_check(hipStreamSynchronize(c->stream), "sync");
        }
        auto t1 = std::chrono::steady_clock::now();
        dirty = true;  // the single-context steps swap u and u1
        return std::chrono::duration<double, std::milli>(t1 - t0).count() / nsteps;
    }
};

extern "C" {

dgx_status dgx_batch_create(int32_t n, const dgx_mesh_desc* meshes, const dgx_space_desc* spaces,
                            const dgx_physics* physics, int32_t precision, dgx_batch** out) {
    if (!out) { g_error = "out is null"; return DGX_E_INVALID; }
    *out = nullptr;
    auto* b = new (std::nothrow) dgx_batch();
    if (!b) { g_error = "out of host memory"; return DGX_E_NOMEM; }
    dgx_status st = guarded([&] {
        if (n <= 0 || !meshes || !spaces || !physics) throw std::invalid_argument("bad batch description");
        if (physics->kind != DGX_EULER) throw Fail(DGX_E_UNSUPPORTED, "batched instances run the Euler residual");
        if (precision == DGX_FP32_BF16VOL)
            throw Fail(DGX_E_UNSUPPORTED, "batched instances run the FP32 and FP64 operators");
        for (int i = 0; i < n; ++i) {
            if (spaces[i].rule != DGX_RULE_2P) throw Fail(DGX_E_UNSUPPORTED, "batched instances use the degree-2p rule");
            dgx_ctx* c = nullptr;
            if (meshes[i].vertices_per_element == 4)
                throw Fail(DGX_E_UNSUPPORTED, "batched instances are triangle meshes");
            dgx_status s = dgx_create(&meshes[i], &spaces[i], physics, precision, &c);
            if (s != DGX_OK) throw Fail(s, std::string("instance ") + std::to_string(i) + ": " + g_error);
            b->inst.push_back(c);
            b->order_mask |= 1 << (c->p - 1);
            if (!b->tabs[c->p - 1]) b->tabs[c->p - 1] = c->tabL.p;
            c->dt = 0.0;
        }
        // identical tables per order need the same basis across instances
        for (auto* c : b->inst)
            if (c->basis != b->inst[0]->basis) throw std::invalid_argument("batched instances must share one basis");
        // matrix-core orders: every instance of the order has the lane tables (contexts of one
        // batch are created alike, so this holds for all or none)
        for (int p = 1; p <= 4; ++p) {
            const void* t = nullptr;
            bool all = true, any = false;
            for (auto* c : b->inst)
                if (c->p == p) {
                    any = true;
                    all = all && c->tabML.p;
                    if (!t) t = c->tabML.p;
                }
            b->tabm[p - 1] = (any && all) ? t : nullptr;
        }
        b->precision = precision;
        int dev = 0;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
        b->cus = prop.multiProcessorCount;
        hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "stream");
        hip_check(hipEventCreate(&b->ev0), "event");
        hip_check(hipEventCreate(&b->ev1), "event");
        hip_check(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "event");
        for (int p = 0; p < 4; ++p) {
            hip_check(hipStreamCreateWithFlags(&b->ostream[p], hipStreamNonBlocking), "stream");
            hip_check(hipEventCreateWithFlags(&b->join[p], hipEventDisableTiming), "event");
        }
    });
    if (st != DGX_OK) { delete b; return st; }
    *out = b;
    return DGX_OK;
}

void dgx_batch_destroy(dgx_batch* b) { delete b; }

dgx_status dgx_batch_instance(dgx_batch* b, int32_t i, dgx_ctx** ctx) {
    return guarded([&] {
        if (!b || !ctx || i < 0 || i >= (int)b->inst.size()) throw std::invalid_argument("bad instance");
        *ctx = b->inst[i];
    });
}

dgx_status dgx_batch_set_dt(dgx_batch* b, const double* dt) {
    return guarded([&] {
        if (!b || !dt) throw std::invalid_argument("null argument");
        for (size_t i = 0; i < b->inst.size(); ++i) {
            if (!(dt[i] > 0) || !std::isfinite(dt[i])) throw std::invalid_argument("dt must be positive");
            b->inst[i]->dt = dt[i];
        }
        b->dirty = true;
    });
}

dgx_status dgx_batch_rk3_steps(dgx_batch* b, int32_t nsteps) {
    return guarded([&] {
        if (!b || nsteps < 0) throw std::invalid_argument("bad arguments");
        for (auto* c : b->inst)
            if (!(c->dt > 0)) throw std::invalid_argument("set the time steps first (dgx_batch_set_dt)");
        // the single-context calls may have swapped state buffers since the last capture
        b->dirty = true;
        b->steps(nsteps);
        hip_check(hipStreamSynchronize(b->stream), "sync");
    });
}

dgx_status dgx_batch_profile(dgx_batch* b, int32_t nsteps, int32_t sequential, double* ms) {
    return guarded([&] {
        if (!b || nsteps <= 0 || !ms) throw std::invalid_argument("bad arguments");
        for (auto* c : b->inst)
            if (!(c->dt > 0)) throw std::invalid_argument("set the time steps first (dgx_batch_set_dt)");
        b->dirty = true;
        auto saved = b->save();
        *ms = sequential ? b->time_sequential(nsteps) : b->time_batched(nsteps);
        b->restore(saved);
        b->dirty = true;
    });
}

dgx_status dgx_batch_get_tuning(const dgx_batch* b, dgx_batch_tuning* t) {
    if (!b || !t) { g_error = "null argument"; return DGX_E_INVALID; }
    *t = b->tune;
    return DGX_OK;
}

dgx_status dgx_batch_set_tuning(dgx_batch* b, const dgx_batch_tuning* t) {
    return guarded([&] {
        if (!b || !t) throw std::invalid_argument("null argument");
        if (t->layout < 0 || t->layout > 2 || t->blocks_per_cu < 1 || t->blocks_per_cu > 16)
            throw std::invalid_argument("bad tuning");
        for (int p = 0; p < 4; ++p)
            if (t->chunk[p] < 1 || t->chunk[p] > 64) throw std::invalid_argument("bad tuning chunk");
        b->tune = *t;
        b->dirty = true;
    });
}

dgx_status dgx_batch_autotune(dgx_batch* b, int32_t trial_steps, dgx_batch_tuning* chosen) {
    return guarded([&] {
        if (!b || trial_steps <= 0) throw std::invalid_argument("bad arguments");
        for (auto* c : b->inst)
            if (!(c->dt > 0)) throw std::invalid_argument("set the time steps first (dgx_batch_set_dt)");
        auto saved = b->save();
        auto measure = [&]() {
            b->dirty = true;
            b->restore(saved);
            return b->time_batched(trial_steps);
        };
        dgx_batch_tuning best = b->tune;
        double best_ms = measure();
        const int chunks[] = {1, 2, 4, 8, 16};
        for (int layout = 0; layout <= 2; ++layout) {
            dgx_batch_tuning cur = best;
            cur.layout = layout;
            for (int p = 4; p >= 1; --p) {
                if (!(b->order_mask & (1 << (p - 1)))) continue;
                for (int ch : chunks) {
                    dgx_batch_tuning t = cur;
                    t.chunk[p - 1] = ch;
                    b->tune = t;
                    double ms = measure();
                    if (ms < best_ms) { best_ms = ms; best = t; cur = t; }
                }
                cur.chunk[p - 1] = best.layout == layout ? best.chunk[p - 1] : cur.chunk[p - 1];
            }
        }
        for (int bpc : {1, 2, 4, 8}) {
            dgx_batch_tuning t = best;
            t.blocks_per_cu = bpc;
            b->tune = t;
            double ms = measure();
            if (ms < best_ms) { best_ms = ms; best = t; }
        }
        b->tune = best;
        b->dirty = true;
        b->restore(saved);
        if (chosen) *chosen = best;
    });
}

}  // extern "C"

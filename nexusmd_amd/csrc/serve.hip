// serve.hip -- continuous batching of many small systems on one GPU.
//
// No counterpart in the reference (its Engine runs one system, start to
// finish, core/baseline/src/Engine.cpp:162-188).  Requests are admitted in
// arrival order while their estimated device footprint fits under the HBM
// budget and the active count is below max_active; every active system lives
// in one Batch and advances in the same batched launches (one hipGraph replay
// per step for the whole batch).  When a request finishes, its state is
// copied out and the batch is re-laid out on the next chunk boundary with the
// survivors (their current state carried over) plus newly admitted requests.
// The mixed path accumulates forces in fixed point and every per-system work
// item is laid out independently of its neighbours, so a system's trajectory
// is bit-identical alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/nexusmd.h"
#include "nx_host.h"

namespace {

struct Request {
    int64_t id = 0;
    nx::HostSystem sys;
    int nsteps = 0, done = 0;
    double dt = 0.0;
    int status = 0;   // 0 queued, 1 running, 2 done, 3 failed
    size_t bytes = 0;
    nx_energies energies{};
    std::string error;
};

size_t estimate_bytes(const nx::HostSystem& h) {
    // atoms, padded slots, pair list and PME grids (FP64 upper bound)
    const double n = h.n;
    const double L[3] = {std::fabs(h.ub[0] - h.lb[0]), std::fabs(h.ub[1] - h.lb[1]), std::fabs(h.ub[2] - h.lb[2])};
    double g[3];
    for (int k = 0; k < 3; ++k) {
        if (h.nb.grid[k] > 0) g[k] = h.nb.grid[k];
        else {
            int m = (int)std::ceil(L[k]) * 10, p = 1;
            while (p < m) p *= 2;
            g[k] = p;
        }
    }
    const double G = g[0] * g[1] * g[2];
    const double slots = n * 1.3 + 512;
    const double list = slots / 64.0 * 800.0 * 8.0;
    return (size_t)(n * 200.0 + slots * 100.0 + list + G * 40.0 + 1.0e6);
}

}  // namespace

struct nx_server {
    int device = 0, precision = 0, max_active = 64;
    int64_t budget = 0, in_use = 0;
    std::unique_ptr<nx::EngineBase> eng;
    std::map<int64_t, std::unique_ptr<Request>> reqs;
    std::deque<int64_t> queue;
    std::vector<int64_t> active;        // order = system index in the batch
    bool dirty = true;
    int64_t next_id = 1;
    int64_t submitted = 0, admitted = 0, finished = 0, failed = 0, batched_steps = 0;
    std::vector<double> step_ms;         // one sample per (request, step)
    double sim_ns = 0.0, wall_s = 0.0;
    std::string err;
    int chunk = 50;
};

namespace {

template <class F>
nx_status sguard(nx_server* s, F&& f) {
    if (!s) return NX_ERR_ARG;
    try {
        f();
        s->err.clear();
        return NX_OK;
    } catch (const nx::HipError& e) {
        s->err = e.what();
        return e.code;
    } catch (const std::invalid_argument& e) {
        s->err = e.what();
        return NX_ERR_ARG;
    } catch (const std::exception& e) {
        s->err = e.what();
        return NX_ERR_HIP;
    }
}

// Copy the device state of every active request back to its host copy.
void sync_active_to_host(nx_server* s) {
    if (!s->eng || s->dirty) return;
    for (size_t i = 0; i < s->active.size(); ++i) {
        Request& r = *s->reqs[s->active[i]];
        s->eng->device_state((int)i, r.sys.pos.data(), r.sys.vel.data());
        r.sys.build_pos.resize(r.sys.pos.size());
        int due = 0;
        s->eng->device_list_state((int)i, r.sys.build_pos.data(), &due);
        r.sys.rebuild_pending = due != 0;
    }
}

void admit(nx_server* s) {
    bool synced = false;
    while (!s->queue.empty() && (int)s->active.size() < s->max_active) {
        Request& r = *s->reqs[s->queue.front()];
        if (s->in_use + (int64_t)r.bytes > s->budget && !s->active.empty()) break;
        // the batch is about to change: keep the running systems' device state first
        if (!synced) { sync_active_to_host(s); synced = true; }
        s->queue.pop_front();
        r.status = 1;
        s->in_use += (int64_t)r.bytes;
        s->active.push_back(r.id);
        s->admitted++;
        s->dirty = true;
    }
}

void relayout(nx_server* s) {
    if (!s->dirty) return;
    if (s->active.empty()) { s->dirty = false; return; }
    if (!s->eng) s->eng.reset(nx::make_engine(s->device, s->precision));
    // A request whose first list build cannot fit fails alone: it leaves the
    // batch and the others are laid out again.
    while (!s->active.empty()) {
        std::vector<nx::HostSystem*> v;
        for (int64_t id : s->active) v.push_back(&s->reqs[id]->sys);
        try {
            s->eng->setup(v);
            break;
        } catch (const nx::HipError& e) {
            const std::vector<int> bad = s->eng->failed_systems();
            if (e.code != NX_ERR_CAPACITY || bad.empty()) throw;
            std::vector<int64_t> keep;
            for (size_t i = 0; i < s->active.size(); ++i) {
                Request& r = *s->reqs[s->active[i]];
                if (std::find(bad.begin(), bad.end(), (int)i) == bad.end()) {
                    keep.push_back(r.id);
                    continue;
                }
                r.status = 3;
                r.error = std::string("admission failed: ") + e.what();
                s->failed++;
                s->in_use -= (int64_t)r.bytes;
            }
            if (keep.size() == s->active.size()) throw;
            s->active = keep;
        }
    }
    s->dirty = false;
}

void retire_finished(nx_server* s) {
    std::vector<int64_t> keep;
    bool any = false;
    for (int64_t id : s->active) any |= (s->reqs[id]->done >= s->reqs[id]->nsteps);
    if (!any) return;
    // energies of the final state for the retiring systems
    std::vector<nx_energies> E;
    s->eng->compute(NX_FORCE_ALL, false, 0.0, false, nullptr, nullptr, nullptr, &E);
    for (size_t i = 0; i < s->active.size(); ++i) {
        Request& r = *s->reqs[s->active[i]];
        if (r.done >= r.nsteps) {
            s->eng->device_state((int)i, r.sys.pos.data(), r.sys.vel.data());
            r.energies = E[i];
            r.status = 2;
            s->finished++;
            s->in_use -= (int64_t)r.bytes;
        } else {
            keep.push_back(r.id);
        }
    }
    sync_active_to_host(s);
    s->active = keep;
    s->dirty = true;
}

}  // namespace

extern "C" {

nx_status nx_server_create(int device, int precision, int64_t budget, int max_active, nx_server** out) {
    if (!out || budget <= 0 || max_active <= 0) return NX_ERR_ARG;
    if (precision != NX_PRECISION_FP64 && precision != NX_PRECISION_MIXED) return NX_ERR_ARG;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return NX_ERR_HIP;
    auto* s = new (std::nothrow) nx_server;
    if (!s) return NX_ERR_NOMEM;
    s->device = device;
    s->precision = precision;
    s->budget = budget;
    s->max_active = max_active;
    *out = s;
    return NX_OK;
}

nx_status nx_server_destroy(nx_server* s) {
    if (!s) return NX_ERR_ARG;
    try { s->eng.reset(); } catch (...) {}
    delete s;
    return NX_OK;
}

const char* nx_server_last_error(const nx_server* s) { return s ? s->err.c_str() : "null server"; }

nx_status nx_submit(nx_server* s, const nx_topology* t, const double lb[3], const double ub[3], const double* pos,
                    const double* vel, int nsteps, double dt, int64_t* req) {
    return sguard(s, [&] {
        if (!t || !lb || !ub || !pos || nsteps < 0 || !(dt > 0)) throw std::invalid_argument("bad submit arguments");
        // reuse the single-context validation through a scratch context
        nx_ctx* c = nullptr;
        nx_status st = nx_create(s->device, s->precision, &c);
        if (st != NX_OK) throw nx::HipError(st, "cannot create a validation context");
        st = nx_set_topology(c, t);
        std::string msg = nx_last_error(c);
        nx_destroy(c);
        if (st != NX_OK) throw std::invalid_argument(msg);
        auto r = std::make_unique<Request>();
        r->id = s->next_id++;
        nx::HostSystem& h = r->sys;
        const int n = t->n_atoms;
        h.n = n;
        h.mass.assign(t->mass, t->mass + n);
        h.q.assign(t->charge, t->charge + n);
        h.sig.assign(t->sigma, t->sigma + n);
        h.eps.assign(t->epsilon, t->epsilon + n);
        h.bonds.assign(t->bonds, t->bonds + t->n_bonds);
        h.angles.assign(t->angles, t->angles + t->n_angles);
        h.torsions.assign(t->torsions, t->torsions + t->n_torsions);
        h.nb = t->nb;
        if (t->excl_offsets && t->excl_indices) {
            h.excl_off.assign(t->excl_offsets, t->excl_offsets + n + 1);
            h.excl_idx.assign(t->excl_indices, t->excl_indices + t->excl_offsets[n]);
        } else {
            nx::exclusions_from_bonds(n, h.bonds, h.excl_off, h.excl_idx);
        }
        for (int k = 0; k < 3; ++k) {
            if (!(std::fabs(ub[k] - lb[k]) > 0)) throw std::invalid_argument("empty box");
            h.lb[k] = lb[k];
            h.ub[k] = ub[k];
        }
        for (int i = 0; i < 3 * n; ++i)
            if (!std::isfinite(pos[i]) || (vel && !std::isfinite(vel[i]))) throw std::invalid_argument("non-finite state");
        h.pos.assign(pos, pos + 3 * (size_t)n);
        if (vel) h.vel.assign(vel, vel + 3 * (size_t)n);
        else h.vel.assign(3 * (size_t)n, 0.0);
        r->nsteps = nsteps;
        r->dt = dt;
        r->bytes = estimate_bytes(h);
        if ((int64_t)r->bytes > s->budget) throw std::invalid_argument("request exceeds the HBM budget");
        if (req) *req = r->id;
        s->queue.push_back(r->id);
        s->reqs[r->id] = std::move(r);
        s->submitted++;
    });
}

nx_status nx_server_run(nx_server* s, int steps) {
    return sguard(s, [&] {
        int budget_steps = steps > 0 ? steps : 0x7fffffff;
        using clock = std::chrono::steady_clock;
        // host time of the last step completion seen (or of this call's start):
        // retirement, admission and re-layout before a chunk count toward the
        // latency of the chunk's first step
        auto t_prev = clock::now();
        while (budget_steps > 0) {
            if (s->eng && !s->dirty) retire_finished(s);
            admit(s);
            if (s->active.empty()) break;
            relayout(s);
            if (s->active.empty()) continue;     // every admitted request failed its layout
            // chunk: no active request may overshoot; one dt per chunk group
            int k = std::min(budget_steps, s->chunk);
            double dt = s->reqs[s->active[0]]->dt;
            for (int64_t id : s->active) {
                const Request& r = *s->reqs[id];
                k = std::min(k, r.nsteps - r.done);
                if (r.dt != dt) throw std::invalid_argument("all active requests must share one dt");
            }
            if (k <= 0) { retire_finished(s); continue; }
            // steps each system had completed before this chunk (device counters)
            std::vector<int> before(s->active.size(), 0);
            {
                const auto& c0 = s->eng->counters();
                for (size_t i = 0; i < before.size() && i < c0.size(); ++i) before[i] = c0[i].steps_done;
            }
            const auto t0 = clock::now();
            std::vector<double> step_ms;
            try {
                s->eng->run_step_times(k, dt, NX_FORCE_ALL, step_ms);
            } catch (const nx::HipError& e) {
                const std::vector<int> bad_idx = s->eng->failed_systems();
                if (e.code != NX_ERR_CAPACITY || bad_idx.empty()) throw;
                for (int f : bad_idx)
                    if (f < 0 || f >= (int)s->active.size()) throw;
                // Isolate every request whose list overflowed: the device froze each
                // at its last completed step.  The others completed the k steps.
                sync_active_to_host(s);
                const auto& c1 = s->eng->counters();
                std::vector<int64_t> keep;
                for (size_t i = 0; i < s->active.size(); ++i) {
                    Request& r = *s->reqs[s->active[i]];
                    if (std::find(bad_idx.begin(), bad_idx.end(), (int)i) == bad_idx.end()) {
                        r.done += k;
                        keep.push_back(r.id);
                        continue;
                    }
                    r.done += std::max(0, c1[i].steps_done - before[i]);
                    r.status = 3;
                    r.error = std::string(e.what()) + " (this request: system " + std::to_string(i) + ", code " +
                              std::to_string(c1[i].overflow) + ")";
                    s->failed++;
                    s->in_use -= (int64_t)r.bytes;
                }
                s->active = keep;
                s->dirty = true;
                t_prev = clock::now();
                continue;
            }
            const auto t1 = clock::now();
            const double stall_ms = std::chrono::duration<double, std::milli>(t0 - t_prev).count();
            s->wall_s += std::chrono::duration<double>(t1 - t_prev).count();
            t_prev = t1;
            s->batched_steps += k;
            // one latency sample per (request, step): the device time between the
            // ends of successive batched steps, the first step of the chunk with
            // the host stall before it
            if (!step_ms.empty()) step_ms[0] += stall_ms;
            for (int64_t id : s->active) {
                Request& r = *s->reqs[id];
                r.done += k;
                s->sim_ns += k * r.dt * 1e-3;
                s->step_ms.insert(s->step_ms.end(), step_ms.begin(), step_ms.end());
            }
            budget_steps -= k;
        }
        if (s->eng && !s->dirty) retire_finished(s);
        admit(s);
    });
}

nx_status nx_poll(nx_server* s, int64_t req, int* status, int* steps_done) {
    return sguard(s, [&] {
        auto it = s->reqs.find(req);
        if (it == s->reqs.end()) throw std::invalid_argument("unknown request");
        if (status) *status = it->second->status;
        if (steps_done) *steps_done = it->second->done;
    });
}

nx_status nx_fetch(nx_server* s, int64_t req, double* pos, double* vel, nx_energies* e) {
    return sguard(s, [&] {
        auto it = s->reqs.find(req);
        if (it == s->reqs.end()) throw std::invalid_argument("unknown request");
        Request& r = *it->second;
        if (r.status == 3) throw nx::HipError(NX_ERR_CAPACITY, "request failed: " + r.error);
        if (r.status != 2) throw nx::HipError(NX_ERR_STATE, "request not finished");
        if (pos) std::memcpy(pos, r.sys.pos.data(), r.sys.pos.size() * sizeof(double));
        if (vel) std::memcpy(vel, r.sys.vel.data(), r.sys.vel.size() * sizeof(double));
        if (e) *e = r.energies;
    });
}

nx_status nx_server_get_stats(nx_server* s, nx_server_stats* st) {
    return sguard(s, [&] {
        if (!st) throw std::invalid_argument("NULL stats");
        std::memset(st, 0, sizeof(*st));
        st->submitted = s->submitted;
        st->admitted = s->admitted;
        st->finished = s->finished;
        st->failed = s->failed;
        st->active = (int64_t)s->active.size();
        st->wall_s = s->wall_s;
        st->batched_steps = s->batched_steps;
        st->bytes_in_use = s->in_use;
        st->bytes_budget = s->budget;
        st->aggregate_ns_per_day = s->wall_s > 0 ? s->sim_ns / s->wall_s * 86400.0 : 0.0;
        if (!s->step_ms.empty()) {
            std::vector<double> v = s->step_ms;
            auto pct = [&](double p) {
                size_t k = (size_t)std::min<double>(v.size() - 1, std::floor(p * (v.size() - 1) + 0.5));
                std::nth_element(v.begin(), v.begin() + k, v.end());
                return v[k];
            };
            st->p50_step_ms = pct(0.50);
            st->p99_step_ms = pct(0.99);
        }
    });
}

}  // extern "C"

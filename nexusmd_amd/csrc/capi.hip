// capi.hip -- the extern "C" entry points of include/nexusmd.h for single
// contexts.  Every call catches C++ exceptions and returns a status code;
// the text of the last error stays on the context (nx_last_error).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/nexusmd.h"
#include "nx_engine.h"
#include "nx_host.h"

struct nx_ctx {
    int device = 0;
    int precision = 0;
    std::unique_ptr<nx::EngineBase> eng;
    nx::HostSystem sys;
    bool have_top = false, have_box = false, have_state = false, dirty = true;
    int timing = NX_TIMING_OFF;
    nx::Tuning tuning;            // launch shapes for the next layout (when tuning_set)
    bool tuning_set = false;
    std::string err;
};

namespace {

// the calling thread's last error of a context-free entry point (nx_last_error(NULL))
thread_local std::string tl_error;

struct ArgError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

template <class F>
nx_status guarded(nx_ctx* c, F&& f) {
    if (!c) return NX_ERR_ARG;
    try {
        f();
        c->err.clear();
        return NX_OK;
    } catch (const ArgError& e) {
        c->err = e.what();
        return NX_ERR_ARG;
    } catch (const nx::HipError& e) {
        c->err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        c->err = "host allocation failed";
        return NX_ERR_NOMEM;
    } catch (const std::exception& e) {
        c->err = e.what();
        return NX_ERR_NUMERIC;
    }
}

}  // namespace

namespace {

void check_topology(const nx_topology* t) {
    if (!t) throw ArgError("topology is NULL");
    const int n = t->n_atoms;
    if (n <= 0) throw ArgError("n_atoms must be positive");
    if (!t->mass || !t->charge || !t->sigma || !t->epsilon) throw ArgError("mass/charge/sigma/epsilon must be given");
    auto in = [n](int i) { return i >= 0 && i < n; };
    if (t->n_bonds < 0 || t->n_angles < 0 || t->n_torsions < 0) throw ArgError("negative term count");
    if ((t->n_bonds && !t->bonds) || (t->n_angles && !t->angles) || (t->n_torsions && !t->torsions))
        throw ArgError("term array is NULL");
    for (int b = 0; b < t->n_bonds; ++b)
        if (!in(t->bonds[b].p1) || !in(t->bonds[b].p2))
            throw ArgError("bond " + std::to_string(b) + ": particle index out of range");
    for (int a = 0; a < t->n_angles; ++a)
        if (!in(t->angles[a].p1) || !in(t->angles[a].p2) || !in(t->angles[a].p3))
            throw ArgError("angle " + std::to_string(a) + ": particle index out of range");
    for (int k = 0; k < t->n_torsions; ++k)
        if (!in(t->torsions[k].p1) || !in(t->torsions[k].p2) || !in(t->torsions[k].p3) || !in(t->torsions[k].p4))
            throw ArgError("torsion " + std::to_string(k) + ": particle index out of range");
    const nx_nonbonded_opts& o = t->nb;
    if (!(o.cutoff > 0)) throw ArgError("cutoff must be positive");
    if (!(o.ewald_tolerance > 0 && o.ewald_tolerance < 0.5)) throw ArgError("ewald_tolerance must be in (0, 0.5)");
    if (o.pme_order != 0 && o.pme_order != 5) throw ArgError("pme_order must be 5");
    if (o.skin < 0 || o.skin > 1.0) throw ArgError("skin must be in [0, 1] nm");
    // the device applies the switch from r_switch >= 0 (a switch distance at or
    // beyond the cutoff leaves every pair unswitched, as in the reference)
    if (o.use_switch && !(std::isfinite(o.switch_distance) && o.switch_distance >= 0))
        throw ArgError("use_switch needs a finite switch_distance >= 0");
    if (t->excl_offsets && t->excl_indices) {
        const int* off = t->excl_offsets;
        if (off[0] != 0) throw ArgError("excl_offsets[0] must be 0");
        for (int i = 0; i < n; ++i)
            if (off[i + 1] < off[i]) throw ArgError("excl_offsets must be non-decreasing");
        for (int e = 0; e < off[n]; ++e)
            if (!in(t->excl_indices[e])) throw ArgError("exclusion index out of range");
    }
    for (int i = 0; i < n; ++i)
        if (!(std::isfinite(t->mass[i]) && t->mass[i] >= 0 && std::isfinite(t->charge[i]) &&
              std::isfinite(t->sigma[i]) && std::isfinite(t->epsilon[i]) && t->epsilon[i] >= 0))
            throw ArgError("non-finite or negative particle parameter at atom " + std::to_string(i));
}

void copy_topology(nx::HostSystem& h, const nx_topology* t) {
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
    if (h.nb.pme_order == 0) h.nb.pme_order = 5;
    if (t->excl_offsets && t->excl_indices) {
        h.excl_off.assign(t->excl_offsets, t->excl_offsets + n + 1);
        h.excl_idx.assign(t->excl_indices, t->excl_indices + t->excl_offsets[n]);
        // keep each row sorted (lookups on the device binary-search nothing, but tests compare sets)
        for (int i = 0; i < n; ++i) std::sort(h.excl_idx.begin() + h.excl_off[i], h.excl_idx.begin() + h.excl_off[i + 1]);
    } else {
        nx::exclusions_from_bonds(n, h.bonds, h.excl_off, h.excl_idx);
    }
    h.list_scale = 1.0;
}

void check_positions(int n, const double* p) {
    for (int i = 0; i < 3 * n; ++i)
        if (!std::isfinite(p[i])) throw ArgError("non-finite position or velocity component " + std::to_string(i));
}

// After a list overflow inside a call the device froze the system at the last
// step it completed.  Its state becomes the host copy and the next call lays
// the context out again with twice the list room, from that state.
template <class F>
void with_capacity_recovery(nx_ctx* c, F&& f) {
    try {
        f();
    } catch (const nx::HipError& e) {
        if (e.code == NX_ERR_CAPACITY && c->eng && !c->dirty) {
            try {
                c->eng->device_state(0, c->sys.pos.data(), c->sys.vel.data());
                c->sys.build_pos.clear();
                c->sys.rebuild_pending = false;
                c->sys.list_scale *= 2.0;
                c->dirty = true;
            } catch (...) {
            }
        }
        throw;
    }
}

void ensure(nx_ctx* c) {
    if (!c->have_top) throw ArgError("nx_set_topology has not been called");
    if (!c->have_box) throw ArgError("nx_set_box has not been called");
    if (!c->have_state) throw ArgError("nx_set_state has not been called");
    if (c->dirty) {
        if (!c->eng) c->eng.reset(nx::make_engine(c->device, c->precision));
        c->eng->set_timing(c->timing);
        if (c->tuning_set) c->eng->set_tuning(c->tuning);
        c->eng->setup({&c->sys});
        c->dirty = false;
    }
}

}  // namespace

extern "C" {

const char* nx_status_string(nx_status s) {
    switch (s) {
        case NX_OK: return "ok";
        case NX_ERR_ARG: return "invalid argument";
        case NX_ERR_HIP: return "HIP runtime error";
        case NX_ERR_CAPACITY: return "capacity exceeded";
        case NX_ERR_STATE: return "call out of order";
        case NX_ERR_NOMEM: return "out of device memory";
        case NX_ERR_NUMERIC: return "non-finite result";
        default: return "unknown status";
    }
}

nx_status nx_create(int device, int precision, nx_ctx** out) {
    if (!out) return NX_ERR_ARG;
    *out = nullptr;
    if (precision != NX_PRECISION_FP64 && precision != NX_PRECISION_MIXED) return NX_ERR_ARG;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return NX_ERR_HIP;
    if (device < 0 || device >= count) return NX_ERR_ARG;
    auto* c = new (std::nothrow) nx_ctx;
    if (!c) return NX_ERR_NOMEM;
    c->device = device;
    c->precision = precision;
    *out = c;
    return NX_OK;
}

nx_status nx_destroy(nx_ctx* c) {
    if (!c) return NX_ERR_ARG;
    try {
        c->eng.reset();
    } catch (...) {
    }
    delete c;
    return NX_OK;
}

const char* nx_last_error(const nx_ctx* c) {
    if (c) return c->err.c_str();
    return tl_error.empty() ? "null context" : tl_error.c_str();
}

nx_status nx_set_topology(nx_ctx* c, const nx_topology* t) {
    return guarded(c, [&] {
        check_topology(t);
        if (c->have_state && c->sys.n != t->n_atoms) c->have_state = false;
        copy_topology(c->sys, t);
        c->have_top = true;
        c->dirty = true;
    });
}

nx_status nx_set_box(nx_ctx* c, const double lb[3], const double ub[3]) {
    return guarded(c, [&] {
        if (!lb || !ub) throw ArgError("box bounds are NULL");
        for (int k = 0; k < 3; ++k) {
            if (!std::isfinite(lb[k]) || !std::isfinite(ub[k]) || !(std::fabs(ub[k] - lb[k]) > 0))
                throw ArgError("box extent must be finite and positive");
            c->sys.lb[k] = lb[k];
            c->sys.ub[k] = ub[k];
        }
        c->have_box = true;
        c->dirty = true;
    });
}

nx_status nx_box_from_positions(int n, const double* pos, double lb[3], double ub[3]) {
    if (n <= 0 || !pos || !lb || !ub) return NX_ERR_ARG;
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) lo[k] = hi[k] = pos[k];
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            const double v = pos[3 * i + k];
            if (!std::isfinite(v)) return NX_ERR_ARG;
            if (v < lo[k]) lo[k] = v;
            if (hi[k] < v) hi[k] = v;
        }
    for (int k = 0; k < 3; ++k) {
        const double margin = 0.01 * std::fabs(hi[k] - lo[k]);
        lb[k] = lo[k] - margin;
        ub[k] = hi[k] + margin;
    }
    return NX_OK;
}

nx_status nx_set_state(nx_ctx* c, const double* pos, const double* vel) {
    return guarded(c, [&] {
        if (!c->have_top) throw ArgError("nx_set_topology must precede nx_set_state");
        if (!pos) throw ArgError("positions are NULL");
        const int n = c->sys.n;
        check_positions(n, pos);
        if (vel) check_positions(n, vel);
        c->sys.pos.assign(pos, pos + 3 * (size_t)n);
        if (vel) c->sys.vel.assign(vel, vel + 3 * (size_t)n);
        else c->sys.vel.assign(3 * (size_t)n, 0.0);
        c->sys.build_pos.clear();          // new positions: the list is built from them
        c->sys.rebuild_pending = false;
        c->have_state = true;
        if (!c->dirty && c->eng) c->eng->set_state(pos, vel);
    });
}

nx_status nx_get_list_state(nx_ctx* c, double* build_pos, int* rebuild_due) {
    return guarded(c, [&] {
        if (c->dirty && c->have_top && c->have_state && !c->sys.build_pos.empty()) {
            // set by nx_set_list_state and not laid out on the device yet
            if (build_pos) std::memcpy(build_pos, c->sys.build_pos.data(), c->sys.build_pos.size() * sizeof(double));
            if (rebuild_due) *rebuild_due = c->sys.rebuild_pending ? 1 : 0;
            return;
        }
        ensure(c);
        std::vector<double> bp(3 * (size_t)c->sys.n);
        int due = 0;
        c->eng->device_list_state(0, bp.data(), &due);
        if (build_pos) std::memcpy(build_pos, bp.data(), bp.size() * sizeof(double));
        if (rebuild_due) *rebuild_due = due;
    });
}

nx_status nx_set_list_state(nx_ctx* c, const double* build_pos, int rebuild_due) {
    return guarded(c, [&] {
        if (!c->have_state) throw ArgError("nx_set_state must precede nx_set_list_state");
        if (!build_pos) throw ArgError("build positions are NULL");
        const int n = c->sys.n;
        check_positions(n, build_pos);
        if (!c->dirty && c->eng) {
            // the device state is the current one: keep it as the host copy the
            // next layout starts from
            c->eng->device_state(0, c->sys.pos.data(), c->sys.vel.data());
        }
        c->sys.build_pos.assign(build_pos, build_pos + 3 * (size_t)n);
        c->sys.rebuild_pending = rebuild_due != 0;
        c->dirty = true;                   // laid out again on the next call, the list built from build_pos
    });
}

nx_status nx_get_state(nx_ctx* c, double* pos, double* vel, double* forces) {
    return guarded(c, [&] {
        if (c->dirty && c->have_top && c->have_state && !forces) {
            // not laid out on the device (yet, or again after a capacity error):
            // the host copy is the current state
            if (pos) std::memcpy(pos, c->sys.pos.data(), c->sys.pos.size() * sizeof(double));
            if (vel) std::memcpy(vel, c->sys.vel.data(), c->sys.vel.size() * sizeof(double));
            return;
        }
        ensure(c);
        std::vector<double> p, v;
        c->eng->get_state(pos ? &p : nullptr, vel ? &v : nullptr);
        if (pos) std::memcpy(pos, p.data(), p.size() * sizeof(double));
        if (vel) std::memcpy(vel, v.data(), v.size() * sizeof(double));
        if (forces) {
            std::vector<double> F;
            c->eng->compute(NX_FORCE_ALL, false, 0.0, false, &F, nullptr, nullptr, nullptr);
            std::memcpy(forces, F.data(), F.size() * sizeof(double));
        }
    });
}

nx_status nx_compute_forces(nx_ctx* c, int mask, double* forces, nx_energies* e) {
    return guarded(c, [&] {
        if (mask & ~NX_FORCE_ALL) throw ArgError("unknown force bits");
        ensure(c);
        std::vector<double> F;
        std::vector<nx_energies> E;
        with_capacity_recovery(c, [&] { c->eng->compute(mask, true, 0.0, false, forces ? &F : nullptr, nullptr, nullptr, &E); });
        if (forces) std::memcpy(forces, F.data(), F.size() * sizeof(double));
        if (e) *e = E[0];
        if (e && !std::isfinite(e->potential)) throw std::runtime_error("non-finite energy");
    });
}

nx_status nx_step(nx_ctx* c, int nsteps, double dt, int mask) {
    return guarded(c, [&] {
        if (nsteps < 0 || !(dt > 0) || !std::isfinite(dt)) throw ArgError("nsteps >= 0 and dt > 0 required");
        if (mask & ~NX_FORCE_ALL) throw ArgError("unknown force bits");
        ensure(c);
        if (nsteps) with_capacity_recovery(c, [&] { c->eng->run(nsteps, dt, mask); });
    });
}

nx_status nx_step_report(nx_ctx* c, double dt, int mask, double* pos, double* vel, double* forces, nx_energies* e) {
    return guarded(c, [&] {
        if (!(dt > 0) || !std::isfinite(dt)) throw ArgError("dt > 0 required");
        if (mask & ~NX_FORCE_ALL) throw ArgError("unknown force bits");
        ensure(c);
        std::vector<double> F, p, v;
        std::vector<nx_energies> E;
        with_capacity_recovery(c, [&] {
            c->eng->compute(mask, true, dt, true, forces ? &F : nullptr, pos ? &p : nullptr, vel ? &v : nullptr, &E);
        });
        if (forces) std::memcpy(forces, F.data(), F.size() * sizeof(double));
        if (pos) std::memcpy(pos, p.data(), p.size() * sizeof(double));
        if (vel) std::memcpy(vel, v.data(), v.size() * sizeof(double));
        if (e) *e = E[0];
    });
}

nx_status nx_step_report_async(nx_ctx* c, double dt, int mask, double* pos, double* vel, double* forces,
                               int64_t* ticket) {
    return guarded(c, [&] {
        if (!(dt > 0) || !std::isfinite(dt)) throw ArgError("dt > 0 required");
        if (mask & ~NX_FORCE_ALL) throw ArgError("unknown force bits");
        if (!ticket) throw ArgError("ticket is NULL");
        ensure(c);
        *ticket = c->eng->report_async(mask, dt, pos, vel, forces);
    });
}

nx_status nx_report_wait(nx_ctx* c, int64_t ticket, nx_energies* e) {
    return guarded(c, [&] {
        if (!c->eng || c->dirty) throw ArgError("no report in flight on this context");
        std::vector<nx_energies> E;
        with_capacity_recovery(c, [&] { c->eng->report_wait(ticket, e ? &E : nullptr); });
        if (e) *e = E[0];
    });
}

nx_status nx_host_alloc(size_t bytes, void** out) {
    if (!out) return NX_ERR_ARG;
    *out = nullptr;
    if (bytes == 0) return NX_OK;
    const hipError_t e = hipHostMalloc(out, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        *out = nullptr;
        tl_error = std::string("hipHostMalloc: ") + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? NX_ERR_NOMEM : NX_ERR_HIP;
    }
    return NX_OK;
}

nx_status nx_host_free(void* p) {
    if (!p) return NX_OK;
    const hipError_t e = hipHostFree(p);
    if (e != hipSuccess) {
        tl_error = std::string("hipHostFree: ") + hipGetErrorString(e);
        return NX_ERR_HIP;
    }
    return NX_OK;
}

nx_status nx_neighbor_pairs(nx_ctx* c, int* pairs, int64_t cap, int64_t* count) {
    return guarded(c, [&] {
        if (cap < 0 || (cap > 0 && !pairs)) throw ArgError("bad pair buffer");
        ensure(c);
        std::vector<int> out;
        long long n = c->eng->export_pairs(cap > 0 ? &out : nullptr);
        if (count) *count = n;
        const long long w = std::min<long long>(n, cap);
        if (w > 0) std::memcpy(pairs, out.data(), sizeof(int) * 2 * (size_t)w);
    });
}

nx_status nx_enable_timing(nx_ctx* c, int on) {
    return guarded(c, [&] {
        if (on < NX_TIMING_OFF || on > NX_TIMING_GRAPH_PAIR) throw ArgError("unknown timing mode");
        c->timing = on;
        if (c->eng) c->eng->set_timing(c->timing);
    });
}

nx_status nx_kernel_time(nx_ctx* c, const char* name, double* ms, int64_t* launches) {
    return guarded(c, [&] {
        if (!name || !ms || !launches) throw ArgError("NULL argument");
        if (!c->eng) throw ArgError("no call has run yet");
        long long n = 0;
        if (!c->eng->kernel_time(name, ms, &n)) { *ms = 0.0; n = 0; }
        *launches = n;
    });
}

nx_status nx_set_tuning(nx_ctx* c, const nx_tuning* t) {
    return guarded(c, [&] {
        if (!t) throw ArgError("NULL tuning");
        if (t->pair_per_cu < 0 || t->pair_per_cu > 16 || t->spread_block < 0 || t->spread_block > 64 ||
            t->fft_xblocks < 0 || t->fft_xblocks > 65536 || t->fft_lines < 0 || t->fft_lines > 8)
            throw ArgError("tuning value out of range");
        nx::Tuning u;
        u.pair_per_cu = t->pair_per_cu;
        u.spread_block = t->spread_block;
        u.fft_xblocks = t->fft_xblocks;
        u.fft_lines = t->fft_lines;
        c->tuning = u;
        c->tuning_set = true;
        if (c->eng && !c->dirty) {
            // keep the device state across the re-layout, with the list built from the same positions
            c->eng->device_state(0, c->sys.pos.data(), c->sys.vel.data());
            c->sys.build_pos.resize(c->sys.pos.size());
            int due = 0;
            c->eng->device_list_state(0, c->sys.build_pos.data(), &due);
            c->sys.rebuild_pending = due != 0;
        }
        c->dirty = true;
    });
}

nx_status nx_get_tuning(nx_ctx* c, nx_tuning* t) {
    return guarded(c, [&] {
        if (!t) throw ArgError("NULL tuning");
        ensure(c);
        const nx::Tuning u = c->eng->tuning();
        t->pair_per_cu = u.pair_per_cu;
        t->spread_block = u.spread_block;
        t->fft_xblocks = u.fft_xblocks;
        t->fft_lines = u.fft_lines;
    });
}

nx_status nx_autotune(nx_ctx* c, int steps, nx_tuning* chosen, double* ms_default, double* ms_best) {
    return guarded(c, [&] {
        if (steps < 4) throw ArgError("steps_per_trial must be >= 4");
        ensure(c);
        const nx::Tuning u = c->eng->autotune(steps, ms_default, ms_best);
        c->tuning = u;
        c->tuning_set = true;
        if (chosen) {
            chosen->pair_per_cu = u.pair_per_cu;
            chosen->spread_block = u.spread_block;
            chosen->fft_xblocks = u.fft_xblocks;
            chosen->fft_lines = u.fft_lines;
        }
    });
}

nx_status nx_get_stats(nx_ctx* c, nx_stats* s) {
    return guarded(c, [&] {
        if (!s) throw ArgError("NULL stats");
        ensure(c);
        c->eng->stats(s);
    });
}

static nx_status fft_common(int device, int precision, int nx_, int ny, int nz, const double* in, double* out,
                            bool forward) {
    if (!in || !out || nx_ < 2 || ny < 2 || nz < 2) return NX_ERR_ARG;
    auto ok = [](int K) {
        int m = K;
        for (int p : {2, 3, 5})
            while (m % p == 0) m /= p;
        return m == 1 && K <= 1024;
    };
    if (!ok(nx_) || !ok(ny) || !ok(nz)) return NX_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return NX_ERR_HIP;
    hipStream_t st;
    if (hipStreamCreate(&st) != hipSuccess) return NX_ERR_HIP;
    if (precision == NX_PRECISION_MIXED) nx::fft3d_roundtrip<nx::PrecMixed>(nx_, ny, nz, in, out, forward, st);
    else nx::fft3d_roundtrip<nx::PrecFP64>(nx_, ny, nz, in, out, forward, st);
    hipError_t e = hipStreamSynchronize(st);
    hipStreamDestroy(st);
    return e == hipSuccess ? NX_OK : NX_ERR_HIP;
}

nx_status nx_fft3d_r2c(int device, int precision, int nx_, int ny, int nz, const double* in, double* out) {
    return fft_common(device, precision, nx_, ny, nz, in, out, true);
}

nx_status nx_fft3d_c2r(int device, int precision, int nx_, int ny, int nz, const double* in, double* out) {
    return fft_common(device, precision, nx_, ny, nz, in, out, false);
}

nx_status nx_fft3d_r2c_noalloc(int device, int precision, int nx_, int ny, int nz, const double* in, double* out) {
    auto set = [](nx_status s, const std::string& m) {
        tl_error = m;
        return s;
    };
    if (!in || !out) return set(NX_ERR_ARG, "nx_fft3d_r2c_noalloc: input or output is NULL");
    if (precision != NX_PRECISION_FP64 && precision != NX_PRECISION_MIXED)
        return set(NX_ERR_ARG, "nx_fft3d_r2c_noalloc: unknown precision");
    auto ok = [](int K) {
        int m = K;
        for (int p : {2, 3, 5})
            while (m % p == 0) m /= p;
        return K >= 2 && m == 1;
    };
    if (!ok(nx_) || !ok(ny) || !ok(nz))
        return set(NX_ERR_ARG, "nx_fft3d_r2c_noalloc: every dimension must be >= 2 with prime factors 2, 3, 5 only");
    std::string err;
    const nx_status s = precision == NX_PRECISION_MIXED
                            ? nx::fft3d_r2c_noalloc<nx::PrecMixed>(device, nx_, ny, nz, in, out, err)
                            : nx::fft3d_r2c_noalloc<nx::PrecFP64>(device, nx_, ny, nz, in, out, err);
    return set(s, s == NX_OK ? std::string() : "nx_fft3d_r2c_noalloc: " + err);
}

}  // extern "C"

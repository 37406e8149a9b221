// dispatcher.cpp -- TaskDispatcher::Simulate on the GPU.
//
// Order of work per step is the reference's Engine::RunSimulation
// (Engine.cpp:162-188): forces at x(t), then the report when the step is a
// report step (step 0 and every step with (step + 1) % interval == 0,
// Engine.cpp:117-158), then the leapfrog update.  Report steps go through
// nx_step_report (forces, energies, kinetic energy with the half-step shift
// and the state before the update); the steps in between run as one nx_step
// call (graph replays on the device, no host round trip per step).
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <exception>
#include <functional>
#include <iostream>
#include <mutex>
#include <stdexcept>
#include <thread>

#include "nexusmd_host.h"

namespace nexusmd {

namespace {

void check(nx_status st, nx_ctx* ctx, const char* what) {
    if (st == NX_OK) return;
    std::string msg = std::string(what) + " failed: " + nx_status_string(st);
    if (ctx) msg += std::string(" (") + nx_last_error(ctx) + ")";
    throw std::runtime_error(msg);
}

struct Ctx {
    nx_ctx* h = nullptr;
    ~Ctx() { if (h) nx_destroy(h); }
};

std::vector<Coords3D> to_coords(const double* a, int n) {
    std::vector<Coords3D> v(static_cast<size_t>(n));
    for (size_t i = 0; i < v.size(); ++i) v[i] = Coords3D{a[3 * i], a[3 * i + 1], a[3 * i + 2]};
    return v;
}
std::vector<Coords3D> to_coords(const std::vector<double>& a) { return to_coords(a.data(), static_cast<int>(a.size() / 3)); }

// The asynchronous report path of Simulate.  Report steps go to the device
// with nx_step_report_async into one of kSets pinned buffer sets; items (a
// report, or the step lines of a stretch of steps) queue in step order.  The
// main thread collects a report (nx_report_wait) once kSets are in flight or
// at a drain, and hands it, with the items before it, to a writer thread that
// formats and writes in order.  A buffer set returns to the pool after the
// writer has used it.  Errors of the writer surface at the next call from the
// main thread.
class ReportPipeline {
   public:
    struct Item {
        bool report = false;
        long long step = 0, to = 0;       // report step, or the stretch [step, to) of echo lines
        nx_energies e{};
        const double *pos = nullptr, *vel = nullptr, *f = nullptr;
        int set = -1;
        int64_t ticket = 0;
    };
    using WriteFn = std::function<void(const Item&)>;
    static constexpr int kSets = 3;

    ReportPipeline(nx_ctx* ctx, int n, WriteFn fn) : ctx_(ctx), n_(n), write_(std::move(fn)) {
        const size_t bytes = 9 * static_cast<size_t>(n) * sizeof(double);
        for (int k = 0; k < kSets; ++k) {
            void* p = nullptr;
            check(nx_host_alloc(bytes, &p), nullptr, "nx_host_alloc");
            sets_[k] = static_cast<double*>(p);
            free_.push_back(k);
        }
        writer_ = std::thread([this] { run(); });
    }
    ~ReportPipeline() {
        stop();
        for (double* p : sets_)
            if (p) nx_host_free(p);
    }

    void report(double dt, int mask, long long step) {
        int k = take_set();
        double* b = sets_[k];
        const size_t m = 3 * static_cast<size_t>(n_);
        Item it;
        it.report = true;
        it.step = step;
        it.set = k;
        it.pos = b;
        it.vel = b + m;
        it.f = b + 2 * m;
        check(nx_step_report_async(ctx_, dt, mask, b, b + m, b + 2 * m, &it.ticket), ctx_, "nx_step_report_async");
        pending_.push_back(it);
        ++in_flight_;
        flush(false);
    }
    void echo(long long from, long long to) {
        Item it;
        it.step = from;
        it.to = to;
        pending_.push_back(it);
        flush(false);
    }
    // every pending item collected and handed to the writer, and the writer idle
    void drain() {
        flush(true);
        std::unique_lock<std::mutex> lk(mu_);
        idle_.wait(lk, [this] { return queue_.empty() && !busy_; });
        rethrow();
    }
    void finish() {
        stop();
        rethrow();
    }

   private:
    int take_set() {
        // a set is free once the writer is done with it; collect reports until one is
        while (true) {
            {
                std::lock_guard<std::mutex> lk(mu_);
                rethrow();
                if (!free_.empty()) {
                    const int k = free_.back();
                    free_.pop_back();
                    return k;
                }
            }
            if (in_flight_ > 0) {
                collect_front();
                continue;
            }
            std::unique_lock<std::mutex> lk(mu_);
            freed_.wait(lk, [this] { return !free_.empty() || error_; });
        }
    }
    // hands the front items to the writer: echo items at once, a report after its copies landed
    void flush(bool all) {
        while (!pending_.empty() && (!pending_.front().report || all || in_flight_ >= kSets)) collect_front();
    }
    void collect_front() {
        Item it = pending_.front();
        pending_.pop_front();
        if (it.report) {
            --in_flight_;
            check(nx_report_wait(ctx_, it.ticket, &it.e), ctx_, "nx_report_wait");
        }
        std::lock_guard<std::mutex> lk(mu_);
        queue_.push_back(it);
        wake_.notify_one();
    }
    void run() {
        while (true) {
            Item it;
            bool failed = false;
            {
                std::unique_lock<std::mutex> lk(mu_);
                busy_ = false;
                idle_.notify_all();
                wake_.wait(lk, [this] { return !queue_.empty() || stop_; });
                if (queue_.empty()) return;
                it = queue_.front();
                queue_.pop_front();
                busy_ = true;
                failed = error_ != nullptr;
            }
            try {
                if (!failed) write_(it);     // after an error the rest is dropped
            } catch (...) {
                std::lock_guard<std::mutex> lk(mu_);
                if (!error_) error_ = std::current_exception();
            }
            if (it.set >= 0) {
                std::lock_guard<std::mutex> lk(mu_);
                free_.push_back(it.set);
                freed_.notify_one();
            }
        }
    }
    void stop() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (stop_) return;
            stop_ = true;
            wake_.notify_all();
        }
        if (writer_.joinable()) writer_.join();
    }
    void rethrow() {
        if (error_) {
            std::exception_ptr e = error_;
            error_ = nullptr;
            std::rethrow_exception(e);
        }
    }

    nx_ctx* ctx_;
    int n_;
    WriteFn write_;
    double* sets_[kSets] = {};
    std::vector<int> free_;
    std::deque<Item> pending_;       // main thread only
    int in_flight_ = 0;              // reports in pending_
    std::deque<Item> queue_;         // to the writer
    std::mutex mu_;
    std::condition_variable wake_, idle_, freed_;
    std::thread writer_;
    bool stop_ = false, busy_ = false;
    std::exception_ptr error_;
};

}  // namespace

std::shared_ptr<TaskDispatcher> TaskDispatcher::CreateDispatcher(const std::string& processorChoice) {
    if (processorChoice == "GPU") return std::make_shared<GpuTaskDispatcher>(NX_PRECISION_FP64);
    if (processorChoice == "GPU_MIXED") return std::make_shared<GpuTaskDispatcher>(NX_PRECISION_MIXED);
    std::cerr << "Invalid processor type. Defaulting to GPU." << std::endl;
    return std::make_shared<GpuTaskDispatcher>(NX_PRECISION_FP64);
}

void TaskDispatcher::initializeParameters() {
    atomPositions.clear();
    masses.clear();
    torsionParams.clear();
    bondParams.clear();
    angleParams.clear();
    nonbondedParams = NonbondedParams{};
}

void GpuTaskDispatcher::Simulate(const std::string& systemFilename, const std::string& stateFilename,
                                 const std::string& inputFilename, const std::string& outputFilename,
                                 double StepSize, int TotalSteps, int interval) {
    initializeParameters();
    if (harmonicBondForceEnabled) std::cout << "Harmonic Bond Force is enabled" << std::endl;
    if (harmonicAngleForceEnabled) std::cout << "Harmonic Angle Force is enabled" << std::endl;
    if (periodicTorsionForceEnabled) std::cout << "Periodic Torsion Force is enabled" << std::endl;
    if (nonbondedForceEnabled) std::cout << "Nonbonded Force is enabled" << std::endl;
    if (!(StepSize > 0) || !std::isfinite(StepSize)) throw std::invalid_argument("StepSize must be positive");
    if (TotalSteps < 0) throw std::invalid_argument("TotalSteps must be >= 0");
    if (interval <= 0) throw std::invalid_argument("interval must be positive");

    lastIngest = IngestTimes{};
    using clock = std::chrono::steady_clock;
    auto since = [](clock::time_point t) { return std::chrono::duration<double>(clock::now() - t).count(); };
    auto t_stage = clock::now();
    SystemDescription sys = SystemXMLParser::Parse(systemFilename);
    lastIngest.parse_system = since(t_stage);
    t_stage = clock::now();
    const bool extended = useStateVelocities || useStateBoxVectors || restart;
    StateData state;
    if (extended) {
        state = StateXMLParser::parseState(stateFilename);
        atomPositions = state.positions;
    } else {
        atomPositions = StateXMLParser::parseStateXML(stateFilename);
    }
    lastIngest.parse_state = since(t_stage);
    masses = sys.masses;
    bondParams = sys.bonds;
    angleParams = sys.angles;
    torsionParams = sys.torsions;
    nonbondedParams = sys.nonbonded;
    const int n = static_cast<int>(atomPositions.size());
    if (n == 0) throw std::runtime_error("no positions in " + stateFilename);
    if (static_cast<int>(masses.size()) != n)
        throw std::runtime_error("system has " + std::to_string(masses.size()) + " particles, state has " +
                                 std::to_string(n) + " positions");

    // per-particle charges and LJ parameters (zero when the system has no NonbondedForce)
    std::vector<double> q(n, 0.0), sig(n, 0.0), eps(n, 0.0);
    if (nonbondedParams.present) {
        if (static_cast<int>(nonbondedParams.particles.size()) != n)
            throw std::runtime_error("NonbondedForce lists " + std::to_string(nonbondedParams.particles.size()) +
                                     " particles for " + std::to_string(n) + " atoms");
        for (int i = 0; i < n; ++i) {
            q[i] = nonbondedParams.particles[i].q;
            sig[i] = nonbondedParams.particles[i].sig;
            eps[i] = nonbondedParams.particles[i].eps;
        }
    } else if (nonbondedForceEnabled) {
        throw std::runtime_error("EnableNonbondedForce: no NonbondedForce in " + systemFilename);
    }

    std::vector<double> pos(3 * static_cast<size_t>(n)), vel(3 * static_cast<size_t>(n), 0.0);   // v(0) = 0 (Initializer.cpp:21)
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) pos[3 * i + k] = atomPositions[i][k];
    if (useStateVelocities) {
        if (!state.hasVelocities || static_cast<int>(state.velocities.size()) != n)
            throw std::runtime_error("EnableStateVelocities: " + stateFilename + " has " +
                                     std::to_string(state.velocities.size()) + " velocities for " + std::to_string(n) +
                                     " positions");
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) vel[3 * i + k] = state.velocities[i][k];
    }

    nx_topology top{};
    top.n_atoms = n;
    top.mass = masses.data();
    top.n_bonds = static_cast<int>(bondParams.size());
    top.bonds = bondParams.data();
    top.n_angles = static_cast<int>(angleParams.size());
    top.angles = angleParams.data();
    top.n_torsions = static_cast<int>(torsionParams.size());
    top.torsions = torsionParams.data();
    top.charge = q.data();
    top.sigma = sig.data();
    top.epsilon = eps.data();
    top.excl_offsets = nullptr;        // 1-2, 1-3, 1-4 from the bonds (Exclusions.cpp:7-44)
    top.excl_indices = nullptr;
    top.nb.cutoff = nonbondedParams.present ? nonbondedParams.cutoff : 1.0;
    top.nb.ewald_tolerance = nonbondedParams.present ? nonbondedParams.ewaldTolerance : 5e-4;
    top.nb.grid[0] = top.nb.grid[1] = top.nb.grid[2] = 0;     // reference grid rule (the XML nx/ny/nz are not used)
    top.nb.pme_order = 5;
    top.nb.use_switch = nonbondedParams.useSwitchingFunction;
    top.nb.switch_distance = nonbondedParams.switchingDistance;
    top.nb.skin = skin;

    Ctx ctx;
    check(nx_create(device, precision_, &ctx.h), nullptr, "nx_create");
    t_stage = clock::now();
    check(nx_set_topology(ctx.h, &top), ctx.h, "nx_set_topology");
    lastIngest.topology = since(t_stage);
    t_stage = clock::now();
    double lb[3], ub[3];
    if (restart && state.hasBoxBounds) {
        // the exact bounds the checkpointed run used
        for (int k = 0; k < 3; ++k) { lb[k] = state.lb[k]; ub[k] = state.ub[k]; }
    } else if (useStateBoxVectors) {
        if (!state.hasBoxVectors) throw std::runtime_error("EnableStateBoxVectors: no PeriodicBoxVectors in " + stateFilename);
        const Coords3D &A = state.boxA, &B = state.boxB, &C = state.boxC;
        if (A[1] != 0 || A[2] != 0 || B[0] != 0 || B[2] != 0 || C[0] != 0 || C[1] != 0)
            throw std::runtime_error("EnableStateBoxVectors: only rectangular boxes are supported");
        lb[0] = lb[1] = lb[2] = 0.0;
        ub[0] = A[0];
        ub[1] = B[1];
        ub[2] = C[2];
    } else {
        // box = bbox of x(0) +- 1% (PeriodicBoundaryCondition.cpp:11-88)
        check(nx_box_from_positions(n, pos.data(), lb, ub), nullptr, "nx_box_from_positions");
    }
    check(nx_set_box(ctx.h, lb, ub), ctx.h, "nx_set_box");
    check(nx_set_state(ctx.h, pos.data(), vel.data()), ctx.h, "nx_set_state");
    long long first = 0;          // number of the first step of this run
    if (restart) {
        first = state.stepCount;
        if (state.hasListState) {
            if (static_cast<int>(state.buildPositions.size()) != n)
                throw std::runtime_error("EnableRestart: neighbour-list state of " + stateFilename + " has " +
                                         std::to_string(state.buildPositions.size()) + " positions for " +
                                         std::to_string(n) + " atoms");
            std::vector<double> bp(3 * static_cast<size_t>(n));
            for (int i = 0; i < n; ++i)
                for (int k = 0; k < 3; ++k) bp[3 * i + k] = state.buildPositions[i][k];
            check(nx_set_list_state(ctx.h, bp.data(), state.rebuildDue), ctx.h, "nx_set_list_state");
        }
    }
    lastIngest.box_state = since(t_stage);
    t_stage = clock::now();
    {
        nx_stats st;      // lays the system out on the device: upload, first list build, PME tables
        check(nx_get_stats(ctx.h, &st), ctx.h, "nx_get_stats");
    }
    lastIngest.device_setup = since(t_stage);
    lastIngest.atoms = n;
    // a restartable state file after `done` steps
    auto checkpoint = [&](long long done) {
        StateData s;
        std::vector<double> p(3 * static_cast<size_t>(n)), v(3 * static_cast<size_t>(n)), bp(3 * static_cast<size_t>(n));
        int due = 0;
        check(nx_get_state(ctx.h, p.data(), v.data(), nullptr), ctx.h, "nx_get_state");
        check(nx_get_list_state(ctx.h, bp.data(), &due), ctx.h, "nx_get_list_state");
        s.positions = to_coords(p);
        s.velocities = to_coords(v);
        s.buildPositions = to_coords(bp);
        s.hasVelocities = s.hasBoxVectors = s.hasBoxBounds = s.hasListState = true;
        s.boxA = Coords3D(ub[0] - lb[0], 0, 0);
        s.boxB = Coords3D(0, ub[1] - lb[1], 0);
        s.boxC = Coords3D(0, 0, ub[2] - lb[2]);
        s.lb = Coords3D(lb[0], lb[1], lb[2]);
        s.ub = Coords3D(ub[0], ub[1], ub[2]);
        s.rebuildDue = due;
        s.stepCount = done;
        s.time = static_cast<double>(done) * StepSize;
        StateXMLParser::writeState(checkpointFilename, s);
    };

    const int mask = (harmonicBondForceEnabled ? NX_FORCE_BOND : 0) | (harmonicAngleForceEnabled ? NX_FORCE_ANGLE : 0) |
                     (periodicTorsionForceEnabled ? NX_FORCE_TORSION : 0) |
                     (nonbondedForceEnabled ? NX_FORCE_NONBONDED : 0);
    auto is_report = [&](long long step) { return step == 0 || ((step + 1) % interval) == 0; };

    const bool ckpt = !checkpointFilename.empty();
    auto is_ckpt = [&](long long step) {
        return ckpt && checkpointInterval > 0 && step > first && step % checkpointInterval == 0;
    };

    Reporter reporter;
    std::vector<PDBAtom> tmpl;
    // a restart appends to the report files: the PDB template without a new header
    if (first > 0) reporter.outputTemplateExtractor(inputFilename, tmpl);
    const long long end = first + TotalSteps;

    // Writes one report (the reference's order: PDB model, TotalEnergy row, PVF
    // block, then the step line) or prints the step lines of a stretch of steps.
    auto write_report = [&](long long step, const nx_energies& e, const double* p, const double* v, const double* fo) {
        const std::vector<Coords3D> P = to_coords(p, n), V = to_coords(v, n), F = to_coords(fo, n);
        const double pe = e.reference_pe, ke = e.kinetic, te = pe + ke;
        const int model = static_cast<int>((step + 1) / interval);
        if (step == 0) reporter.pdbOutputGeneratorPart1(inputFilename, outputFilename, tmpl);
        reporter.pdbOutputGeneratorPart2(outputFilename, tmpl, P, model);
        reporter.TotalEnergyReport(outputFilename, ke, pe, te, static_cast<int>(step));
        reporter.TestPVFReport(outputFilename, P, V, F, static_cast<int>(step));
        if (echo_steps) std::cout << "Simulation Step #" << step + 1 << " is completed" << std::endl;
    };
    auto echo = [&](long long from, long long to) {
        if (echo_steps)
            for (long long s = from; s < to; ++s) std::cout << "Simulation Step #" << s + 1 << " is completed\n";
    };

    if (!async_reports) {
        std::vector<double> f(3 * static_cast<size_t>(n));
        long long step = first;
        while (step < end) {
            if (is_ckpt(step)) checkpoint(step);
            if (is_report(step)) {
                nx_energies e{};
                check(nx_step_report(ctx.h, StepSize, mask, pos.data(), vel.data(), f.data(), &e), ctx.h,
                      "nx_step_report");
                write_report(step, e, pos.data(), vel.data(), f.data());
                ++step;
                continue;
            }
            // the steps up to the next report step or checkpoint run as one device call
            long long next = step + 1;
            while (next < end && !is_report(next) && !is_ckpt(next)) ++next;
            check(nx_step(ctx.h, static_cast<int>(next - step), StepSize, mask), ctx.h, "nx_step");
            echo(step, next);
            step = next;
        }
    } else {
        // Report steps are enqueued without waiting (nx_step_report_async) into
        // pinned buffer sets; a writer thread formats and writes each report once
        // its copies have landed, in step order, while the device runs ahead.
        ReportPipeline pipe(ctx.h, n, [&](const ReportPipeline::Item& it) {
            if (it.report) write_report(it.step, it.e, it.pos, it.vel, it.f);
            else echo(it.step, it.to);
        });
        long long step = first;
        while (step < end) {
            if (is_ckpt(step)) {
                pipe.drain();
                checkpoint(step);
            }
            if (is_report(step)) {
                pipe.report(StepSize, mask, step);
                ++step;
                continue;
            }
            long long next = step + 1;
            while (next < end && !is_report(next) && !is_ckpt(next)) ++next;
            check(nx_step(ctx.h, static_cast<int>(next - step), StepSize, mask), ctx.h, "nx_step");
            pipe.echo(step, next);
            step = next;
        }
        pipe.drain();
        pipe.finish();
    }
    if (ckpt) checkpoint(end);
    std::cout.flush();
}

}  // namespace nexusmd

// dispatcher.cpp -- TaskDispatcher::Simulate on the GPU.
//
// Order of work per step is the reference's Engine::RunSimulation
// (Engine.cpp:162-188): forces at x(t), then the report when the step is a
// report step (step 0 and every step with (step + 1) % interval == 0,
// Engine.cpp:117-158), then the leapfrog update.  Report steps go through
// nx_step_report (forces, energies, kinetic energy with the half-step shift
// and the state before the update); the steps in between run as one nx_step
// call (graph replays on the device, no host round trip per step).
#include <cmath>
#include <iostream>
#include <stdexcept>

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

std::vector<Coords3D> to_coords(const std::vector<double>& a) {
    std::vector<Coords3D> v(a.size() / 3);
    for (size_t i = 0; i < v.size(); ++i) v[i] = Coords3D{a[3 * i], a[3 * i + 1], a[3 * i + 2]};
    return v;
}

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

    SystemDescription sys = SystemXMLParser::Parse(systemFilename);
    atomPositions = StateXMLParser::parseStateXML(stateFilename);
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
    check(nx_set_topology(ctx.h, &top), ctx.h, "nx_set_topology");
    double lb[3], ub[3];
    check(nx_box_from_positions(n, pos.data(), lb, ub), nullptr, "nx_box_from_positions");   // box = bbox of x(0) +- 1% (PeriodicBoundaryCondition.cpp:11-88)
    check(nx_set_box(ctx.h, lb, ub), ctx.h, "nx_set_box");
    check(nx_set_state(ctx.h, pos.data(), vel.data()), ctx.h, "nx_set_state");

    const int mask = (harmonicBondForceEnabled ? NX_FORCE_BOND : 0) | (harmonicAngleForceEnabled ? NX_FORCE_ANGLE : 0) |
                     (periodicTorsionForceEnabled ? NX_FORCE_TORSION : 0) |
                     (nonbondedForceEnabled ? NX_FORCE_NONBONDED : 0);
    auto is_report = [&](int step) { return step == 0 || ((step + 1) % interval) == 0; };

    Reporter reporter;
    std::vector<PDBAtom> tmpl;
    std::vector<double> f(3 * static_cast<size_t>(n));
    int step = 0;
    while (step < TotalSteps) {
        if (is_report(step)) {
            nx_energies e{};
            check(nx_step_report(ctx.h, StepSize, mask, pos.data(), vel.data(), f.data(), &e), ctx.h,
                  "nx_step_report");
            const auto P = to_coords(pos), V = to_coords(vel), F = to_coords(f);
            const double pe = e.reference_pe, ke = e.kinetic, te = pe + ke;
            const int model = (step + 1) / interval;
            if (step == 0) reporter.pdbOutputGeneratorPart1(inputFilename, outputFilename, tmpl);
            reporter.pdbOutputGeneratorPart2(outputFilename, tmpl, P, model);
            reporter.TotalEnergyReport(outputFilename, ke, pe, te, step);
            reporter.TestPVFReport(outputFilename, P, V, F, step);
            if (echo_steps) std::cout << "Simulation Step #" << step + 1 << " is completed" << std::endl;
            ++step;
            continue;
        }
        int next = step;
        while (next < TotalSteps && !is_report(next)) ++next;
        check(nx_step(ctx.h, next - step, StepSize, mask), ctx.h, "nx_step");
        if (echo_steps)
            for (int s = step; s < next; ++s) std::cout << "Simulation Step #" << s + 1 << " is completed\n";
        step = next;
    }
    std::cout.flush();
}

}  // namespace nexusmd

// nx_host.h -- host-side system description and the engine interface used by
// the C-ABI (single contexts and the batching server).
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nexusmd.h"
#include "nx_device.h"

namespace nx {

// Engine error carrying the C-ABI status it maps to.
struct HipError : std::runtime_error {
    nx_status code;
    HipError(nx_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// One system as copied in through the C-ABI.
struct HostSystem {
    int n = 0;
    std::vector<double> mass, q, sig, eps;
    std::vector<nx_bond> bonds;
    std::vector<nx_angle> angles;
    std::vector<nx_torsion> torsions;
    std::vector<int> excl_off, excl_idx;     // CSR, local indices
    nx_nonbonded_opts nb{};
    double lb[3] = {0, 0, 0}, ub[3] = {0, 0, 0};
    std::vector<double> pos, vel;            // AoS [n][3]
    double list_scale = 1.0;                 // pair-list capacity multiplier
    // Neighbour-list state carried across a re-layout of a batch: the positions
    // the current list was built from (AoS; empty = build from pos) and whether
    // a rebuild was already due.  Rebuilding from the same positions gives the
    // same list, so a system's trajectory does not depend on when the batch
    // around it changes.
    std::vector<double> build_pos;
    bool rebuild_pending = false;
};

void exclusions_from_bonds(int n, const std::vector<nx_bond>& bonds, std::vector<int>& off, std::vector<int>& idx);

// Launch shapes the autotuner chooses among (0 = the built-in rule).  None of
// them changes what a kernel computes for a given atom, pair or grid line, so
// every choice gives the same bits in the mixed path.
struct Tuning {
    int pair_per_cu = 0;     // persistent pair-kernel blocks per CU (built-in: min(occupancy, 3))
    int spread_block = 0;    // spread tile: columns per block side (built-in: the cost model's choice)
    int fft_xblocks = 0;     // plane FFT: x-pass blocks the kz range is split into (built-in: ~512)
    int fft_lines = 0;       // line FFT: lines per block (built-in: as many as fit 60 KiB of LDS, <= 8)
};

class EngineBase {
   public:
    virtual ~EngineBase() = default;
    virtual void setup(const std::vector<HostSystem*>& systems) = 0;
    virtual void run(int nsteps, double dt, int mask) = 0;
    // as run, with the device time of every step (ms, from the end of the previous step)
    virtual void run_step_times(int nsteps, double dt, int mask, std::vector<double>& ms) = 0;
    virtual void compute(int mask, bool kinetic, double dt, bool integrate, std::vector<double>* F,
                         std::vector<double>* pos, std::vector<double>* vel, std::vector<nx_energies>* E) = 0;
    virtual void get_state(std::vector<double>* pos, std::vector<double>* vel) = 0;
    // one report step enqueued without waiting (see nx_step_report_async); returns its ticket
    virtual long long report_async(int mask, double dt, double* pos, double* vel, double* F) = 0;
    // waits for a ticket's copies; energies per system
    virtual void report_wait(long long ticket, std::vector<nx_energies>* E) = 0;
    virtual void set_state(const double* pos, const double* vel) = 0;
    virtual long long export_pairs(std::vector<int>* out) = 0;   // nullptr: count only
    virtual void stats(nx_stats* st) = 0;
    virtual bool kernel_time(const std::string& name, double* ms, long long* n) = 0;
    virtual void set_timing(int mode) = 0;   // NX_TIMING_*
    virtual void set_graphs(bool on) = 0;
    virtual int nsys() const = 0;
    virtual const std::vector<int>& failed_systems() const = 0;   // overflowed in the last call
    virtual const std::vector<SysCount>& counters() const = 0;
    virtual hipStream_t stream() const = 0;
    virtual void device_state(int s, double* pos, double* vel) = 0;
    // build positions (AoS) of system s's current neighbour list and its rebuild flag
    virtual void device_list_state(int s, double* build_pos, int* rebuild_due) = 0;
    virtual size_t bytes() const = 0;
    // launch shapes: set (a re-layout follows at the next setup), read, or chosen by timed trials
    virtual void set_tuning(const Tuning& t) = 0;
    virtual Tuning tuning() const = 0;
    // Times `steps` steps per candidate, knob by knob, on the current systems and
    // their current state (restored after every trial and at the end), keeps the
    // fastest, and caches it under the batch's key.  Returns ms/step of the
    // built-in shapes and of the choice.
    virtual Tuning autotune(int steps, double* ms_default, double* ms_best) = 0;
};

EngineBase* make_engine(int device, int precision);

}  // namespace nx

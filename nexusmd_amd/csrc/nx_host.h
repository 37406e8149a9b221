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
};

EngineBase* make_engine(int device, int precision);

}  // namespace nx

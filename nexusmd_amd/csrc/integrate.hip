// integrate.hip -- per-step packing, leapfrog update, kinetic energy.
//
// Leapfrog as VerletIntegration::Advance (core/baseline/src/
// VerletIntegration.cpp:49-89): v += F/m dt; x += v dt; one wrap into
// [lb, ub]; atoms with 1/m = 0 are left untouched.  Kinetic energy with the
// half-step shift of KineticEnergy::calculateKineticEnergy
// (core/baseline/src/KineticEnergy.cpp:21-41).  Positions and velocities stay
// FP64 in both precision paths; the mixed path converts its fixed-point
// force sums to FP64 here.
#include "nx_device.h"
#include "nx_engine.h"

namespace nx {

// Slot-order position (x, y, z, q) for the pair and PME kernels.
template <class P>
__device__ __forceinline__ typename P::real4 slot_position(double x, double y, double z, double q) {
    using real = typename P::real;
    typename P::real4 p;
    p.x = (real)x;
    p.y = (real)y;
    p.z = (real)z;
    p.w = (real)q;
    return p;
}

// Slot-order copy of the positions for the pair and PME kernels; clears F_slot.
// Needed only where positions or slot forces changed outside a step (the
// list build and the update keep both current otherwise).
template <class P>
__global__ void pack_kernel(const SysDev* __restrict__ sd, const SysCount* __restrict__ sc,
                            const int* __restrict__ slot_sys, const int* __restrict__ atom_of_slot,
                            const double* __restrict__ X, const double* __restrict__ Y, const double* __restrict__ Z,
                            const double* __restrict__ XB, const double* __restrict__ YB,
                            const double* __restrict__ ZB, const double* __restrict__ Q, typename P::real4* pos,
                            typename P::facc* FX, typename P::facc* FY, typename P::facc* FZ, int n_slots_total) {
    using real = typename P::real;
    const int sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl >= n_slots_total) return;
    const int s = slot_sys[sl];
    const int a = (sl - sd[s].slot_off < sc[s].n_slots) ? atom_of_slot[sl] : -1;
    typename P::real4 p;
    if (a >= 0) {
        double c[3] = {X[a], Y[a], Z[a]};
        if constexpr (P::mixed) continuous_position(sd[s], X[a], Y[a], Z[a], XB[a], YB[a], ZB[a], c);
        p = slot_position<P>(c[0], c[1], c[2], Q[a]);
    } else {
        p.x = p.y = p.z = (real)kPadPos;
        p.w = 0;
    }
    pos[sl] = p;
    FX[sl] = typename P::facc{};
    FY[sl] = typename P::facc{};
    FZ[sl] = typename P::facc{};
}

template <class P>
__device__ __forceinline__ void total_force(int a, int sl, const typename P::facc* FAx, const typename P::facc* FAy,
                                            const typename P::facc* FAz, const typename P::facc* FSx,
                                            const typename P::facc* FSy, const typename P::facc* FSz, double* f) {
    if constexpr (P::mixed) {
        // exact integer sum, then one conversion
        long long x = FAx[a] + FSx[sl], y = FAy[a] + FSy[sl], z = FAz[a] + FSz[sl];
        f[0] = (double)x * (1.0 / kForceScale);
        f[1] = (double)y * (1.0 / kForceScale);
        f[2] = (double)z * (1.0 / kForceScale);
    } else {
        // exact integer sums of the two parts, then one conversion
        f[0] = fix2_value(FAx[a].hi + FSx[sl].hi, FAx[a].lo + FSx[sl].lo);
        f[1] = fix2_value(FAy[a].hi + FSy[sl].hi, FAy[a].lo + FSy[sl].lo);
        f[2] = fix2_value(FAz[a].hi + FSz[sl].hi, FAz[a].lo + FSz[sl].lo);
    }
}

template <class P>
__global__ void kinetic_kernel(const int* __restrict__ atom_sys, const int* __restrict__ slot_of_atom,
                               const double* __restrict__ VX, const double* __restrict__ VY,
                               const double* __restrict__ VZ, const double* __restrict__ mass,
                               const typename P::facc* FAx, const typename P::facc* FAy, const typename P::facc* FAz,
                               const typename P::facc* FSx, const typename P::facc* FSy, const typename P::facc* FSz,
                               double dt, long long* energy, int n_atoms) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    double ke = 0.0;
    int s = -1;
    if (a < n_atoms) {
        s = atom_sys[a];
        double f[3];
        total_force<P>(a, slot_of_atom[a], FAx, FAy, FAz, FSx, FSy, FSz, f);
        double v[3] = {VX[a], VY[a], VZ[a]};
        const double m = mass[a];
        if (m > 0) {
            const double im = 1.0 / m, ts = 0.5 * dt;
            v[0] += f[0] * im * ts;
            v[1] += f[1] * im * ts;
            v[2] += f[2] * im * ts;
        }
        ke = 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    wave_energy_add(energy, s < 0 ? 0 : s, E_KIN, ke, s >= 0);
}

// The largest squared displacement of the block's atoms, into slot
// blockIdx % kStepMaxSlots of step_max (one atomic per block, few blocks per
// slot; every thread calls it).  The next nl_waves launch reads and clears
// the slots and reports the maximum to the host.
__device__ __forceinline__ void block_max_displacement(unsigned* step_max, float dm2) {
    __shared__ float wmax[16];
    float m = dm2;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if (lane == 0) wmax[wv] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < nw; ++w) m = fmaxf(m, wmax[w]);
        // non-negative floats order as their bit patterns
        atomicMax(&step_max[blockIdx.x % kStepMaxSlots], __float_as_uint(m));
    }
}

// Leapfrog update; also writes the slot-order position the next step's force
// kernels read (x, y, z, q at the atom's slot) and clears the slot force it
// has consumed, so no packing pass runs between builds.
template <class P>
__global__ void integrate_kernel(const SysDev* __restrict__ sd, SysCount* sc, const int* __restrict__ atom_sys,
                                 const int* __restrict__ slot_of_atom, double* X, double* Y, double* Z, double* VX,
                                 double* VY, double* VZ, const double* __restrict__ invm,
                                 const double* __restrict__ Q, const double* __restrict__ XB,
                                 const double* __restrict__ YB, const double* __restrict__ ZB,
                                 typename P::facc* FAx, typename P::facc* FAy, typename P::facc* FAz,
                                 typename P::facc* FSx, typename P::facc* FSy, typename P::facc* FSz,
                                 typename P::real4* pos, double dt, int n_atoms, unsigned* step_max) {
    using real = typename P::real;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    bool moved = false;
    int s = 0;
    float dm2 = 0.0f;       // squared displacement from the build position
    if (a < n_atoms) {
        s = atom_sys[a];
        const SysDev& d = sd[s];
        const int sl = slot_of_atom[a];
        double f[3];
        total_force<P>(a, sl, FAx, FAy, FAz, FSx, FSy, FSz, f);
        FSx[sl] = typename P::facc{};
        FSy[sl] = typename P::facc{};
        FSz[sl] = typename P::facc{};
        // a system whose neighbour-list build overflowed lacks pair forces: it is
        // frozen at the last step it completed (the host reports it as failed)
        const bool frozen = sc[s].overflow != 0;
        if (a == d.atom_off && !frozen) sc[s].steps_done += 1;
        const double im = invm[a];
        double x = X[a], y = Y[a], z = Z[a];
        if (im != 0.0 && !frozen) {
            double vx = VX[a] + im * f[0] * dt;
            double vy = VY[a] + im * f[1] * dt;
            double vz = VZ[a] + im * f[2] * dt;
            double p[3] = {x + vx * dt, y + vy * dt, z + vz * dt};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if (p[k] < d.lb[k]) p[k] += d.L[k];
                else if (p[k] > d.ub[k]) p[k] -= d.L[k];
            }
            x = p[0]; y = p[1]; z = p[2];
            X[a] = x; Y[a] = y; Z[a] = z;
            VX[a] = vx; VY[a] = vy; VZ[a] = vz;
        }
        const double dx = min_image1(x - XB[a], d.L[0]);
        const double dy = min_image1(y - YB[a], d.L[1]);
        const double dz = min_image1(z - ZB[a], d.L[2]);
        if constexpr (P::mixed)
            pos[sl] = slot_position<P>(XB[a] + dx, YB[a] + dy, ZB[a] + dz, Q[a]);     // continuous_position
        else
            pos[sl] = slot_position<P>(x, y, z, Q[a]);
        const double d2 = dx * dx + dy * dy + dz * dz;
        moved = !(d2 <= d.skin_half2);
        dm2 = (float)d2;
    }
    // lanes of a wave storing to one flag merge into one write
    if (moved) sc[s].rebuild = 1;
    if (step_max) block_max_displacement(step_max, dm2);
}

template <class P>
__global__ void collect_kernel(const int* __restrict__ slot_of_atom, const typename P::facc* FAx,
                               const typename P::facc* FAy, const typename P::facc* FAz, const typename P::facc* FSx,
                               const typename P::facc* FSy, const typename P::facc* FSz, double* FX, double* FY,
                               double* FZ, int n_atoms) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n_atoms) return;
    double f[3];
    total_force<P>(a, slot_of_atom[a], FAx, FAy, FAz, FSx, FSy, FSz, f);
    FX[a] = f[0];
    FY[a] = f[1];
    FZ[a] = f[2];
}

// A report's snapshot in atom order, AoS, as three blocks [n][3]: positions,
// velocities, total forces (the layout the reports and the caller take).
template <class P>
__global__ void report_snapshot_kernel(const int* __restrict__ slot_of_atom, const double* __restrict__ X,
                                       const double* __restrict__ Y, const double* __restrict__ Z,
                                       const double* __restrict__ VX, const double* __restrict__ VY,
                                       const double* __restrict__ VZ, const typename P::facc* FAx,
                                       const typename P::facc* FAy, const typename P::facc* FAz,
                                       const typename P::facc* FSx, const typename P::facc* FSy,
                                       const typename P::facc* FSz, double* out, int n_atoms) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n_atoms) return;
    double f[3];
    total_force<P>(a, slot_of_atom[a], FAx, FAy, FAz, FSx, FSy, FSz, f);
    const size_t n = (size_t)n_atoms, i = 3 * (size_t)a;
    out[i] = X[a];
    out[i + 1] = Y[a];
    out[i + 2] = Z[a];
    out[3 * n + i] = VX[a];
    out[3 * n + i + 1] = VY[a];
    out[3 * n + i + 2] = VZ[a];
    out[6 * n + i] = f[0];
    out[6 * n + i + 1] = f[1];
    out[6 * n + i + 2] = f[2];
}

template <class P>
void launch_report_snapshot(Batch<P>& B, double* out) {
    const int T = 256;
    report_snapshot_kernel<P><<<(unsigned)((B.n_atoms + T - 1) / T), T, 0, B.stream>>>(
        B.slot_of_atom, B.X, B.Y, B.Z, B.VX, B.VY, B.VZ, B.FAx, B.FAy, B.FAz, B.FSx, B.FSy, B.FSz, out, B.n_atoms);
}
template void launch_report_snapshot<PrecFP64>(Batch<PrecFP64>&, double*);
template void launch_report_snapshot<PrecMixed>(Batch<PrecMixed>&, double*);

template <class P>
void launch_pack(Batch<P>& B) {
    const int T = 256;
    pack_kernel<P><<<(unsigned)((B.n_slots_cap + T - 1) / T), T, 0, B.stream>>>(
        B.d_sys, B.d_cnt, B.slot_sys, B.atom_of_slot, B.X, B.Y, B.Z, B.XB, B.YB, B.ZB, B.q, B.pos, B.FSx, B.FSy,
        B.FSz, B.n_slots_cap);
}

template <class P>
void launch_kinetic(Batch<P>& B, double dt) {
    const int T = 256;
    kinetic_kernel<P><<<(unsigned)((B.n_atoms + T - 1) / T), T, 0, B.stream>>>(
        B.atom_sys, B.slot_of_atom, B.VX, B.VY, B.VZ, B.mass, B.FAx, B.FAy, B.FAz, B.FSx, B.FSy, B.FSz, dt, B.energy,
        B.n_atoms);
}

template <class P>
void launch_integrate(Batch<P>& B, double dt) {
    const int T = 256;
    integrate_kernel<P><<<(unsigned)((B.n_atoms + T - 1) / T), T, 0, B.stream>>>(
        B.d_sys, B.d_cnt, B.atom_sys, B.slot_of_atom, B.X, B.Y, B.Z, B.VX, B.VY, B.VZ, B.invm, B.q, B.XB, B.YB, B.ZB,
        B.FAx, B.FAy, B.FAz, B.FSx, B.FSy, B.FSz, B.pos, dt, B.n_atoms, B.step_max);
}

template <class P>
void launch_collect_forces(Batch<P>& B, double* FX, double* FY, double* FZ) {
    const int T = 256;
    collect_kernel<P><<<(unsigned)((B.n_atoms + T - 1) / T), T, 0, B.stream>>>(
        B.slot_of_atom, B.FAx, B.FAy, B.FAz, B.FSx, B.FSy, B.FSz, FX, FY, FZ, B.n_atoms);
}

template void launch_pack<PrecFP64>(Batch<PrecFP64>&);
template void launch_pack<PrecMixed>(Batch<PrecMixed>&);
template void launch_kinetic<PrecFP64>(Batch<PrecFP64>&, double);
template void launch_kinetic<PrecMixed>(Batch<PrecMixed>&, double);
template void launch_integrate<PrecFP64>(Batch<PrecFP64>&, double);
template void launch_integrate<PrecMixed>(Batch<PrecMixed>&, double);
template void launch_collect_forces<PrecFP64>(Batch<PrecFP64>&, double*, double*, double*);
template void launch_collect_forces<PrecMixed>(Batch<PrecMixed>&, double*, double*, double*);

}  // namespace nx

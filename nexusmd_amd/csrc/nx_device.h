// nx_device.h -- device-side types and helpers shared by the HIP kernels.
//
// Data layout (all systems of a batch packed back to back):
//   atom order  : FP64 master state x/y/z, vx/vy/vz, 1/m (SoA), per-atom
//                 force accumulators F_atom (bonded terms, exclusion
//                 correction), atom_sys, slot_of_atom.
//   slot order  : atoms binned into xy columns, sorted by z inside a column
//                 and padded to 8-atom clusters; 8 consecutive clusters of a
//                 column form a super-cluster (one wavefront of the pair
//                 kernel).  Per slot: packed (x,y,z,q), (sig/2, 2*sqrt(eps)),
//                 atom_of_slot (-1 = padding), force accumulators F_slot
//                 (pair, PME).
//   PME         : per system a real grid [Kx][Ky][Kz] (fixed-point int64 in
//                 the mixed path, FP64 otherwise), a half-spectrum complex
//                 grid [Kx][Ky][Kz/2+1] and a real potential grid.
// Forces: FP64 path adds doubles with hardware atomics; mixed path converts
// each contribution to 64-bit fixed point (2^32 units per kJ/mol/nm), so sums
// are exact integers and independent of order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nx {

constexpr double kOne4PiEps0 = 138.93545764446428693393;   // kJ nm / (mol e^2)
constexpr double kPi = 3.14159265358979323846;
constexpr double kSqrtPi = 1.77245385090551602729;
constexpr double kTwoOverSqrtPi = 1.1283791670955126;
constexpr double kForceScale = 4294967296.0;               // 2^32 fixed-point units
constexpr double kEnergyScale = 1073741824.0;              // 2^30 fixed-point units
constexpr int kClusterSize = 8;
constexpr int kSuperClusters = 8;                           // clusters per super-cluster
constexpr int kPmeOrder = 5;

constexpr int NX_FORCE_NONBONDED_BIT = 8;   // the nonbonded bit of a force mask

enum EnergyTerm { E_BOND = 0, E_ANGLE, E_TORSION, E_COUL, E_LJ, E_RECIP, E_EXCL, E_KIN, E_NTERMS };

// Per-system constants (device copy; written once per topology/box).
struct SysDev {
    double L[3], invL[3], lb[3], ub[3];
    double rc, rc2, rlist, alpha, skin_half2;
    double sw_on, sw_inv;        // switching: r_switch, 1/(rc - r_switch); sw_on < 0 = off
    int n_atoms, atom_off;
    int slot_off, slot_cap;
    int col_off, ncx, ncy, col_cap;   // columns: ncx*ncy
    int super_off, super_cap;
    int chunk_off, chunk_cap;    // pair-kernel work items (list chunks)
    int list_off, list_cap;
    int emask_off, emask_cap;
    double colw[2];              // column widths
    int grid[3];
    int pair_image;              // mixed path: per-entry periodic images in the list (0: pair by pair)
    long long rgrid_off, cgrid_off;   // offsets into the real / complex grids
    int excl_off, n_excl;        // excluded pair list
    int bond_off, n_bonds, angle_off, n_angles, tors_off, n_tors;
    int tw_off[3];               // twiddle-table offsets per axis
    int sp_block[3];             // spread tiling: columns per block in x, y; z slabs (0 = direct atomics)
    double sp_margin;            // distance an atom may move between rebuilds
};

// Per-system counters written by the device.
struct SysCount {
    int n_slots, n_clusters, n_super, list_count;
    int emask_count, rebuild, overflow, rebuilds;
    int n_chunks, steps_done, pad1, pad2;   // steps_done: leapfrog steps the system completed
};
// overflow bit of a system that ran a step without the list kernels although
// its list was due for a build: the update froze it for that step (the host
// runs the step again with the list kernels; not a failure)
constexpr int kStaleList = 8;

// What nl_waves of a one-system batch reports to host memory early in every
// step, for the host's choice between the step graphs with and without the
// list kernels: the report count, the largest displacement of an atom from
// its build position after the last update, the steps done, whether this
// step's list is due, and the overflow flags.
struct __align__(16) StepInfo {
    unsigned seq;               // reports so far (modulo 2^32)
    float dmax;                 // nm
    int steps_done;
    int flags;                  // bit 0: the list is due; bits 1..: the overflow flags
};
// slots of the update's largest squared displacement (block b adds to slot b % 64)
constexpr int kStepMaxSlots = 64;

constexpr int kChunkEntries = 32;       // pair-list entries per chunk, the pair kernel's unit of work
// Pair-list entries one super-cluster may hold (the list build keeps them in
// LDS).  A super-cluster is 64 atoms of one column, about 3.5 nm tall at water
// density, with ~700-950 j clusters within rlist = 1.1 nm: 1024 was exceeded
// by a few super-clusters of a 1M-atom water box.
constexpr int kSuperListCap = 2048;
constexpr int kPairWaves = 4;           // independent wavefronts per 256-thread pair-kernel block
// Estimated pair-kernel work in quarter tiles: per tile, per list entry (the
// j loads, the j-force reduction and its LDS staging) and per chunk (i staging
// and the force flush).  Fitted to per-wavefront times of the 100k-atom step:
// 0.16 us per tile and 0.46 us per entry.  The list build writes the sum per
// chunk and cuts the chunks into contiguous ranges of equal sums, one per
// wavefront; the pair kernel pools the ranges of each XCD.
constexpr int kCostTile = 4, kCostEntry = 11, kCostChunk = 24;
// The pair kernel's chunk counters, one per XCD, 128 bytes apart.
constexpr int kPoolStride = 32, kPoolInts = 8 * kPoolStride;

// Second word of a pair-list entry: the exclusion-mask id in the low 26 bits
// (kNoMask: the tile has no listed exclusion) and the j-cluster's periodic
// image in the high 6 bits, two per axis (0: none, 1: +L, 2: -L), for which
// every atom pair of the entry that can come within the cutoff before the next
// build is at its nearest image; kImgPerPair (all six bits set): the entry's
// pairs need the nearest image pair by pair.
constexpr unsigned kNoMask = 0x3ffffffu;
constexpr unsigned kImgShift = 26u;
constexpr unsigned kImgPerPair = 0x3fu;

// Slot position of a padding slot.  Finite, so that the pair kernel's force
// products stay finite, and far outside every box, so that a padding slot
// fails every cutoff test; the pair kernel moves padding j slots to -kPadPos
// so that two padding slots are far apart as well.
constexpr double kPadPos = 1e15;

// Issue priority of a wavefront of the kernels that run beside the pair kernel
// (PME, bonded terms), above the pair kernel's default 0: the sequencer
// otherwise issues for the oldest wavefront first, and the pair kernel's
// persistent wavefronts are older than every kernel launched after it.
__device__ __forceinline__ void side_priority() { __builtin_amdgcn_s_setprio(2); }

// ------------------------------------------------------------ precision

// FP64 force accumulator: v = hi 2^-20 + lo 2^-74, each part summed exactly
// in a 64-bit integer, so a sum of contributions is the same whatever their
// order (the reference sums serially and is reproducible; double atomics are
// not).  Range |v| < 8.7e12, resolution 2^-74 (5e-23) per contribution.
struct Fix2 {
    long long hi, lo;
};

struct PrecFP64 {
    using real = double;
    using real2 = double2;
    using real4 = double4;
    using facc = Fix2;            // force accumulator element
    using gacc = long long;       // PME charge grid element, 2^48 fixed point
    static constexpr double grid_scale = 281474976710656.0;
    static constexpr bool mixed = false;
    static constexpr int pair_waves_per_simd = 1;   // minimum pair-kernel waves per SIMD the registers must allow
};

struct PrecMixed {
    using real = float;
    using real2 = float2;
    using real4 = float4;
    using facc = long long;       // 2^32 fixed point
    using gacc = long long;       // PME charge grid element, 2^32 fixed point
    static constexpr double grid_scale = 4294967296.0;
    static constexpr bool mixed = true;
    static constexpr int pair_waves_per_simd = 4;
    // Three pair-kernel waves per SIMD (the persistent grid's three blocks per
    // CU) at 112 VGPRs leave 176 of a SIMD's 512 to the PME kernels beside
    // them: a 1024-thread plane-FFT block (four waves per SIMD at 40 VGPRs)
    // then fits; at 120 it does not, and PME waits for the pair kernel.
    static constexpr int pair_max_vgpr = 112;
};

__device__ __forceinline__ long long to_fixed(float v) {
    float s = v * 4294967296.0f;
    s = fminf(fmaxf(s, -9.2e18f), 9.2e18f);
    return __float2ll_rn(s);
}
__device__ __forceinline__ long long to_fixed(double v) {
    double s = v * kForceScale;
    s = fmin(fmax(s, -9.2e18), 9.2e18);
    return __double2ll_rn(s);
}

// the two parts of v: hi = v 2^20 rounded, lo = the rest times 2^74 rounded
// (v - hi 2^-20 is exact: the low bits of v)
__device__ __forceinline__ void fix2_parts(double v, long long& hi, long long& lo) {
    double s = v * 1048576.0;
    s = fmin(fmax(s, -9.2e18), 9.2e18);
    const double h = rint(s);
    hi = (long long)h;
    lo = (long long)rint((s - h) * 18014398509481984.0);     // (v - hi 2^-20) 2^74 = (s - h) 2^54
}
__device__ __forceinline__ void facc_add(Fix2* p, double v) {
    long long hi, lo;
    fix2_parts(v, hi, lo);
    atomicAdd(reinterpret_cast<unsigned long long*>(&p->hi), static_cast<unsigned long long>(hi));
    atomicAdd(reinterpret_cast<unsigned long long*>(&p->lo), static_cast<unsigned long long>(lo));
}
__device__ __forceinline__ void facc_add(long long* p, float v) {
    atomicAdd(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(to_fixed(v)));
}
__device__ __forceinline__ void facc_add(long long* p, double v) {
    atomicAdd(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(to_fixed(v)));
}
// Non-atomic add / store, for a location that no other thread writes meanwhile.
__device__ __forceinline__ void facc_set_add(Fix2* p, double v) {
    long long hi, lo;
    fix2_parts(v, hi, lo);
    p->hi += hi;
    p->lo += lo;
}
__device__ __forceinline__ void facc_set_add(long long* p, float v) { *p += to_fixed(v); }
__device__ __forceinline__ void facc_set(Fix2* p, double v) {
    long long hi, lo;
    fix2_parts(v, hi, lo);
    p->hi = hi;
    p->lo = lo;
}
__device__ __forceinline__ void facc_set(long long* p, float v) { *p = to_fixed(v); }
__device__ __forceinline__ double fix2_value(long long hi, long long lo) {
    return (double)hi * (1.0 / 1048576.0) + (double)lo * (1.0 / 18889465931478580854784.0);   // 2^-20, 2^-74
}
__device__ __forceinline__ double facc_get(const Fix2* p) { return fix2_value(p->hi, p->lo); }
__device__ __forceinline__ double facc_get(const long long* p) { return static_cast<double>(*p) * (1.0 / kForceScale); }

// Energy accumulation in 2^30 fixed point.  Every contribution is rounded to
// an integer on its own lane before any summation, so a system's energy is the
// same integer whatever the grouping of its terms into waves and blocks (alone
// or inside any batch).
__device__ __forceinline__ long long energy_fixed(double v) {
    double s = v * kEnergyScale;
    s = fmin(fmax(s, -9.2e18), 9.2e18);
    return __double2ll_rn(s);
}
__device__ __forceinline__ void energy_add_fixed(long long* e, long long v) {
    atomicAdd(reinterpret_cast<unsigned long long*>(e), static_cast<unsigned long long>(v));
}

// ---------------------------------------------------------- wave helpers

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ long long wave_sum(long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide (<=1024 threads) sum of fixed-point energies into slot e (one atomic per block).
__device__ __forceinline__ void block_energy_add(long long* e, long long v) {
    __shared__ long long part[16];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) part[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; ++i) t += part[i];
        energy_add_fixed(e, t);
    }
    __syncthreads();
}

// Adds v of the valid lanes to energy[s * E_NTERMS + term] with one atomic per
// wavefront when all valid lanes belong to one system.  Every lane must call it.
__device__ __forceinline__ void wave_energy_add(long long* energy, int s, int term, double v, bool valid) {
    const unsigned long long m = __ballot(valid);
    if (m == 0ull) return;
    const long long q = valid ? energy_fixed(v) : 0ll;
    const int lead = __ffsll((long long)m) - 1;
    const int s0 = __shfl(s, lead, 64);
    if (__all(!valid || s == s0)) {
        const long long t = wave_sum(q);
        if ((int)(threadIdx.x & 63) == lead) energy_add_fixed(&energy[s0 * E_NTERMS + term], t);
    } else if (valid) {
        energy_add_fixed(&energy[s * E_NTERMS + term], q);
    }
}

// Reference minimum image: one +-L correction (PeriodicBoundaryCondition.cpp:93-104),
// written with selects so that it compiles without branches.
__device__ __forceinline__ double min_image1(double d, double L) {
    const double h = 0.5 * L;
    const double up = d - L, dn = d + L;
    return d > h ? up : (d < -h ? dn : d);
}
__device__ __forceinline__ float min_image1(float d, float L) {
    const float h = 0.5f * L;
    const float up = d - L, dn = d + L;
    return d > h ? up : (d < -h ? dn : d);
}

// The slot position the mixed path's pair and PME kernels read: the build-time
// wrapped position plus the displacement since the build, continuous between
// builds (an atom's wrapped position jumps by L when it crosses the box), so
// that a j-cluster's periodic image found at the build holds until the next
// one.  (The FP64 path reads the positions themselves and takes the nearest
// image pair by pair with the reference's arithmetic.)
__device__ __forceinline__ void continuous_position(const SysDev& d, double x, double y, double z, double xb,
                                                    double yb, double zb, double* p) {
    p[0] = xb + min_image1(x - xb, d.L[0]);
    p[1] = yb + min_image1(y - yb, d.L[1]);
    p[2] = zb + min_image1(z - zb, d.L[2]);
}

// r^2 without contraction, left-to-right as Coords3D::dot (Coords3D.h:116-124).
__device__ __forceinline__ double dot_rn(double x, double y, double z) {
    return __dadd_rn(__dadd_rn(__dmul_rn(x, x), __dmul_rn(y, y)), __dmul_rn(z, z));
}

// Cross-lane sums on a 64-wide wavefront without LDS: DPP row rotate for
// lane^8, permlane16/32 swaps for lane^16 and lane^32, DPP quad permutes and
// the half-row mirror for lane^1, lane^2 and the 8-lane halves.
__device__ __forceinline__ int dpp_ror8(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, false); }
__device__ __forceinline__ int dpp_xor1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false); }
__device__ __forceinline__ int dpp_xor2(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false); }
__device__ __forceinline__ int dpp_hmirror(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false); }

// sum over lanes {l, l^8, l^16, ..., l^56} (lanes sharing l & 7)
__device__ __forceinline__ float sum_stride8(float v) {
    v += __int_as_float(dpp_ror8(__float_as_int(v)));
    auto p = __builtin_amdgcn_permlane16_swap(__float_as_int(v), __float_as_int(v), false, false);
    v = __int_as_float(p[0]) + __int_as_float(p[1]);
    auto q = __builtin_amdgcn_permlane32_swap(__float_as_int(v), __float_as_int(v), false, false);
    return __int_as_float(q[0]) + __int_as_float(q[1]);
}
__device__ __forceinline__ double sum_stride8(double v) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    v += __hiloint2double(dpp_ror8(hi), dpp_ror8(lo));
    lo = __double2loint(v); hi = __double2hiint(v);
    auto pl = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    auto ph = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    v = __hiloint2double(ph[0], pl[0]) + __hiloint2double(ph[1], pl[1]);
    lo = __double2loint(v); hi = __double2hiint(v);
    auto ql = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    auto qh = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    return __hiloint2double(qh[0], ql[0]) + __hiloint2double(qh[1], ql[1]);
}
// sum over the 8 lanes l & ~7 .. l | 7
__device__ __forceinline__ float sum_group8(float v) {
    v += __int_as_float(dpp_xor1(__float_as_int(v)));
    v += __int_as_float(dpp_xor2(__float_as_int(v)));
    v += __int_as_float(dpp_hmirror(__float_as_int(v)));
    return v;
}
__device__ __forceinline__ double sum_group8(double v) {
    v += __hiloint2double(dpp_xor1(__double2hiint(v)), dpp_xor1(__double2loint(v)));
    v += __hiloint2double(dpp_xor2(__double2hiint(v)), dpp_xor2(__double2loint(v)));
    v += __hiloint2double(dpp_hmirror(__double2hiint(v)), dpp_hmirror(__double2loint(v)));
    return v;
}

// XCD-aware block order: consecutive work items land on one XCD (8 XCDs,
// round-robin dispatch), so neighbouring super-clusters share an L2.
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
    const int per = (nblocks + 7) >> 3;
    const int xcd = b & 7, k = b >> 3;
    const int full = nblocks & 7 ? (nblocks & 7) : 8;   // XCDs that get `per` blocks
    int base = (xcd < full) ? xcd * per : full * per + (xcd - full) * (per - 1);
    int r = base + k;
    return r < nblocks ? r : b;
}

}  // namespace nx

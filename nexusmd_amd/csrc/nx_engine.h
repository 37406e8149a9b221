// nx_engine.h -- host-side batch of systems resident on one GPU.
//
// A Batch holds every device buffer for one or more independent systems
// packed back to back (a single-system context is a batch of one).  Kernels
// take the packed arrays plus per-item system tables, so one launch advances
// every system of the batch.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/nexusmd.h"

#include "nx_device.h"

namespace nx {

struct HostSystem;   // engine.hip

// Plane FFT passes: threads per block, LDS available to one workgroup on gfx950.
constexpr int kPlaneThreads = 1024;
constexpr int kXPassThreads = 512;
constexpr size_t kMaxBlockLds = 160 * 1024 - 1024;   // less the kernels' static LDS

// LDS-tiled spreading: the tile region holds at most kSpreadLdsPoints grid
// points; its x planes are spread_plane(ny, nz) points apart, an odd count, so
// that the five lanes of one atom (one x plane each) fall on five different
// LDS banks (with an even stride they could all land on one).
constexpr int kSpreadLdsPoints = 7680;   // 60 KB of int64 / double
__host__ __device__ constexpr int spread_plane(int ny, int nz) { return (ny * nz) | 1; }

// LDS bytes of a plane pass (two buffers, the y and z twiddles).
template <typename R2>
inline size_t plane_lds_bytes(int Ky, int Kz) {
    const int np = (Ky + 1) / 2, Kzh = Kz / 2 + 1;
    const int cap = np * Kz > Ky * Kzh ? np * Kz : Ky * Kzh;
    return (size_t)(Ky + Kz + 2 * cap) * sizeof(R2);
}

// LDS bytes of an x pass over nk wave numbers kz (two buffers, the twiddles).
template <typename R2>
inline size_t xpass_lds_bytes(int Kx, int nk) {
    return (size_t)(Kx + 2 * nk * Kx) * sizeof(R2);
}

template <class P>
struct Batch {
    using real = typename P::real;
    using real2 = typename P::real2;
    using real4 = typename P::real4;
    using facc = typename P::facc;

    int nsys = 0;
    std::vector<SysDev> sys;             // host copy of the descriptors
    SysDev* d_sys = nullptr;
    SysCount* d_cnt = nullptr;
    hipStream_t stream = nullptr;

    int n_atoms = 0, n_slots_cap = 0, n_cols = 0, n_super_cap = 0;
    long long n_list_cap = 0, n_emask_cap = 0;

    // atom order (FP64 master state)
    double *X = nullptr, *Y = nullptr, *Z = nullptr;
    double *VX = nullptr, *VY = nullptr, *VZ = nullptr;
    double *invm = nullptr, *mass = nullptr, *q = nullptr, *sig = nullptr, *eps = nullptr;
    double *XB = nullptr, *YB = nullptr, *ZB = nullptr;
    int *atom_sys = nullptr, *slot_of_atom = nullptr, *atom_col = nullptr;
    facc *FAx = nullptr, *FAy = nullptr, *FAz = nullptr;

    // slot order
    int *slot_sys = nullptr, *atom_of_slot = nullptr, *slot_tmp = nullptr;
    real4* pos = nullptr;
    real2* lj = nullptr;
    facc *FSx = nullptr, *FSy = nullptr, *FSz = nullptr;

    // columns / clusters / pair list
    int *col_sys = nullptr, *col_count = nullptr, *col_fill = nullptr, *col_slot = nullptr, *col_super = nullptr;
    int* col_tmp = nullptr;         // bin counts during a build (zero between builds)
    float4 *cl_ctr = nullptr, *cl_half = nullptr;
    float4* slot_wpos = nullptr;    // build-time wrapped position of each slot, w = 1 (atom) or 0 (padding)
    int* super_sys = nullptr;
    int4* super_item = nullptr;
    int* chunk_sys = nullptr;
    int4* chunks = nullptr;         // (first cluster, n clusters, first entry, end entry)
    int* chunk_prefix = nullptr;    // [nsys + 1] running total of chunks over systems
    int* chunk_cost = nullptr;      // per chunk: estimated pair-kernel work (list build)
    int* wave_start = nullptr;      // [pair_waves + 1] chunk ranges of equal estimated work (per XCD: its waves' union)
    int* pair_pool = nullptr;       // [kPoolInts] per XCD: chunks claimed so far by the running pair kernel
    // one-system batches: nl_waves' report to the host (host memory, mapped),
    // the report count (device) and the update's largest squared displacement
    // (device, kStepMaxSlots float bit patterns)
    StepInfo* step_info = nullptr;
    unsigned* step_seq = nullptr;
    unsigned* step_max = nullptr;
    int pair_blocks = 0, pair_waves = 0;
    int n_cu = 0;                   // compute units of the device
    int n_chunk_cap = 0;
    uint2* list = nullptr;
    unsigned long long* emask = nullptr;
    int *excl_off = nullptr, *excl_idx = nullptr;
    int2* excl_pairs = nullptr;
    int n_excl_pairs = 0;

    // bonded terms (global atom indices)
    int2* bond_idx = nullptr;
    double2* bond_par = nullptr;
    int n_bonds = 0;
    int4* angle_idx = nullptr;
    double2* angle_par = nullptr;
    int n_angles = 0;
    real* tF = nullptr;             // per-term forces [3][n_tslots]: bonds 2, angles 3, torsions 4 slots each
    int n_tslots = 0;
    int* at_off = nullptr;          // per atom: CSR of the term-force slots that act on it
    int* at_idx = nullptr;
    int4* tors_idx = nullptr;
    double4* tors_par = nullptr;   // k, phase, periodicity, unused
    int n_tors = 0;

    // PME
    using gacc = typename P::gacc;
    gacc* rgrid = nullptr;          // spread charges [sum Kx*Ky*Kz], fixed point (P::grid_scale)
    real2* cgrid = nullptr;         // half spectrum [sum Kx*Ky*(Kz/2+1)]
    real* phi = nullptr;            // potential grid [sum Kx*Ky*Kz]
    real* eta = nullptr;            // influence function on the half spectrum
    real2* twiddle = nullptr;       // per system, per axis exp(-2 pi i t / K)
    long long rgrid_total = 0, cgrid_total = 0, tw_total = 0;
    int4* spread_tiles = nullptr;   // (sys, column block x, column block y, z slab)
    int n_spread_tiles = 0, n_spread_direct_slots = 0;
    float* slot_zkey = nullptr;     // build-time z of each slot (sorted within a column)
    int4* fftz_items = nullptr;     // (sys, first line, n lines, 0)
    int4* ffty_items = nullptr;     // (sys, x, kz0, 0)
    int4* fftx_items = nullptr;     // (sys, y, kz0, 0)
    int n_fftz = 0, n_ffty = 0, n_fftx = 0;
    int fft_lines = 1, fft_tile = 1, fft_maxK = 0;
    int fft_plane = 0;              // 1: plane passes (fftz_items = planes, fftx_items = x blocks)
    size_t fft_plane_lds = 0, fft_x_lds = 0;
    int pme_any = 0;

    // energies [nsys][E_NTERMS] in 2^30 fixed point
    long long* energy = nullptr;

    // flags
    int timing = 0;                 // NX_TIMING_* mode
};

// part 0: the whole list work; 1: all but its last kernel; 2: the last kernel
// alone, which then also does the pair kernel's set-up (launch_pair_begin)
template <class P> void launch_nlist(Batch<P>& B, bool clear_flags, int part = 0);
// persistent pair-kernel grid: per_cu_request blocks per CU (0: the built-in count), capped by occupancy
template <class P> int pair_grid_blocks(int n_cu, int n_chunk_cap, int per_cu_request);
template <class P> int pair_blocks_occupancy();     // pair-kernel blocks one CU holds at once
template <class P> void launch_pack(Batch<P>& B);
template <class P> void launch_pair(Batch<P>& B, bool energy);
// the two launches of launch_pair: the per-system set-up kernel, then the pair kernel
template <class P> void launch_pair_begin(Batch<P>& B);
template <class P> void launch_pair_kernel(Batch<P>& B, bool energy);
template <class P> void launch_bonded(Batch<P>& B, int mask, bool energy);
template <class P> void launch_atom_forces(Batch<P>& B, int mask, bool energy);
template <class P> void launch_pme_spread(Batch<P>& B);
bool fft_plane_enabled();
template <class P> hipError_t prepare_fft(Batch<P>& B);
template <class P> void launch_fft_forward(Batch<P>& B);
template <class P> void launch_fft_convolve(Batch<P>& B, bool energy);
template <class P> void launch_fft_inverse(Batch<P>& B);
template <class P> void launch_pme_gather(Batch<P>& B);
template <class P> void launch_kinetic(Batch<P>& B, double dt);
template <class P> void launch_integrate(Batch<P>& B, double dt);
template <class P> void launch_collect_forces(Batch<P>& B, double* FX, double* FY, double* FZ);
// positions, velocities and total forces in atom order into out[3][n][3]
template <class P> void launch_report_snapshot(Batch<P>& B, double* out);
// Writes up to cap pairs to out and their total count to *d_count, which must be zero before the launch.
template <class P> void launch_export_pairs(Batch<P>& B, int2* out, long long cap, long long* d_count);
// standalone FFT (verification)
template <class P> void fft3d_roundtrip(int nx, int ny, int nz, const double* in, double* out, bool forward,
                                        hipStream_t st);
// the forward transform of fft3d_roundtrip in the library's static device
// scratch array (no allocation); err receives the reason of a failure
template <class P> nx_status fft3d_r2c_noalloc(int device, int nx, int ny, int nz, const double* in, double* out,
                                               std::string& err);

}  // namespace nx

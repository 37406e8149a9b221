// pme.hip -- smooth PME reciprocal space on the device.
//
// Follows NonbondedForce::ReciprocalPMEcalculateForcesAndEnergy
// (core/baseline/src/Forces/NonbondedForce.cpp:467-499): fractional index
// and offset from the origin (:51-71), order-5 B-splines by the standard
// recursion (:150-181), charge spreading (:220-260), forward 3D FFT, the
// influence-function convolution and energy (:323-388), backward FFT and the
// force gather (:407-461).
//
// The reference runs a full complex-to-complex FFTW transform of a real grid.
// Here the spread grid is real, so the forward transform is real-to-complex
// onto the half spectrum [Kx][Ky][Kz/2+1] and the backward one complex-to-
// real; by Hermitian symmetry this is the same result (the reference keeps
// only the real part after its backward transform).  Five passes, each one
// read and one write of the grid:
//   Z  : R2C along z (lines contiguous), fixed-point -> real on load;
//   Y  : C2C along y on tiles of consecutive kz (coalesced rows);
//   X  : C2C along x, multiply by the influence function eta(k) and sum the
//        energy, inverse C2C along x -- fused so the spectrum makes one trip;
//   Y' : inverse C2C along y;
//   Z' : C2R along z into the real potential grid.
// Each line FFT is a Stockham autosort radix-{4,2,3,5} transform in LDS with
// twiddles from a per-axis table exp(-2 pi i t / K) computed on the host in FP64.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "nx_device.h"
#include "nx_engine.h"

namespace nx {

template <typename real>
__device__ __forceinline__ void bspline5(real dr, real* th, real* dth) {
    // NonbondedForce.cpp:150-181 with pmeOrder = 5
    th[4] = 0;
    th[1] = dr;
    th[0] = 1 - dr;
#pragma unroll
    for (int k = 3; k < 5; ++k) {
        const real div = (real)1.0 / (real)(k - 1.0);
        th[k - 1] = div * dr * th[k - 2];
#pragma unroll
        for (int l = 1; l < (k - 1); ++l)
            th[k - l - 1] = div * ((dr + l) * th[k - l - 2] + (k - l - dr) * th[k - l - 1]);
        th[0] = div * (1 - dr) * th[0];
    }
    dth[0] = -th[0];
#pragma unroll
    for (int k = 1; k < 5; ++k) dth[k] = th[k - 1] - th[k];
    const real div = (real)1.0 / (real)4.0;
    th[4] = div * dr * th[3];
#pragma unroll
    for (int l = 1; l < 4; ++l) th[4 - l] = div * ((dr + l) * th[4 - l - 1] + (5 - l - dr) * th[4 - l]);
    th[0] = div * (1 - dr) * th[0];
}

template <typename real>
__device__ __forceinline__ void grid_index(real x, double invL, int K, int& idx, real& fr) {
    real t = x * (real)invL;
    t = (t - floor(t)) * (real)K;
    if (!(t >= (real)0)) t = 0;
    int ti = (int)t;
    fr = t - (real)ti;
    idx = ti >= K ? ti - K : ti;     // t is in [0, K]
}

// The charge grid is fixed point in both paths (exact, order-free sums): 2^32
// in the mixed path (FP32 contributions), 2^48 in the FP64 path.  A grid
// value is a sum of q w over the atoms within 2.5 grid spacings, a few units
// at most, far inside the FP64 range of +-32768; the resolution, 3.6e-15 per
// contribution, keeps the reciprocal forces within ~1e-12 of their double
// sums (at 2^40 positive and negative charges cancelling on a point left
// 2e-10 relative RMS in the total force).
__device__ __forceinline__ long long grid_fixed(float v) { return to_fixed(v); }
__device__ __forceinline__ long long grid_fixed(double v) {
    double s = v * PrecFP64::grid_scale;
    s = fmin(fmax(s, -9.2e18), 9.2e18);
    return (long long)rint(s);
}
template <typename R>
__device__ __forceinline__ void grid_add(long long* g, R v) {
    atomicAdd(reinterpret_cast<unsigned long long*>(g), static_cast<unsigned long long>(grid_fixed(v)));
}
// LDS accumulation (a ds_add_u64)
template <typename R>
__device__ __forceinline__ void grid_add_lds(long long* g, R v) {
    atomicAdd(reinterpret_cast<unsigned long long*>(g), static_cast<unsigned long long>(grid_fixed(v)));
}
// adding an already accumulated value
__device__ __forceinline__ void grid_add_acc(long long* g, long long v) {
    atomicAdd(reinterpret_cast<unsigned long long*>(g), static_cast<unsigned long long>(v));
}

// One thread per (slot, ix): 25 grid updates each.
template <class P>
__global__ void pme_spread_kernel(const SysDev* __restrict__ sd, const SysCount* __restrict__ sc,
                                  const int* __restrict__ slot_sys, const int* __restrict__ atom_of_slot,
                                  const typename P::real4* __restrict__ pos, typename P::gacc* rgrid,
                                  int n_slots_total) {
    side_priority();
    using real = typename P::real;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int slot = (int)(t / 5), ix = (int)(t % 5);
    if (slot >= n_slots_total) return;
    const int s = slot_sys[slot];
    const SysDev& d = sd[s];
    if (slot - d.slot_off >= sc[s].n_slots || sc[s].overflow) return;
    if (atom_of_slot[slot] < 0) return;
    if (d.grid[0] == 0 || d.sp_block[0] > 0) return;
    const typename P::real4 p = pos[slot];
    const int Kx = d.grid[0], Ky = d.grid[1], Kz = d.grid[2];
    int i0, i1, i2;
    real f0, f1, f2;
    grid_index<real>(p.x, d.invL[0], Kx, i0, f0);
    grid_index<real>(p.y, d.invL[1], Ky, i1, f1);
    grid_index<real>(p.z, d.invL[2], Kz, i2, f2);
    real tx[5], dtx[5], ty[5], dty[5], tz[5], dtz[5];
    bspline5<real>(f0, tx, dtx);
    bspline5<real>(f1, ty, dty);
    bspline5<real>(f2, tz, dtz);
    const real qx = p.w * tx[ix];
    const int xi = (i0 + ix) % Kx;
    typename P::gacc* g = rgrid + d.rgrid_off + (long long)xi * Ky * Kz;
#pragma unroll
    for (int iy = 0; iy < 5; ++iy) {
        const int yi = (i1 + iy) % Ky;
        const real qxy = qx * ty[iy];
#pragma unroll
        for (int iz = 0; iz < 5; ++iz) {
            const int zi = (i2 + iz) % Kz;
            grid_add(&g[(long long)yi * Kz + zi], qxy * tz[iz]);
        }
    }
}

// Quotient of non-negative integers n, d with n + d < 2^22 through a float
// reciprocal: (n + 0.5) / d is at least 0.5 / d away from an integer, and the
// two roundings (of 1/d and of the product) move it by at most
// (n / d + 1) 2^-23, which is less.
__device__ __forceinline__ int qdiv(int n, float inv_d) { return (int)(((float)n + 0.5f) * inv_d); }

// LDS-tiled spreading.  One workgroup per (system, column block, z slab):
// the atoms binned into those columns at the last neighbour-list build whose
// build-time z lies in the slab.  Their stencils are accumulated in an LDS copy
// of the grid region they can reach (block + slab, widened by the distance an
// atom may move between rebuilds, + 4 points for the order-5 stencil), then
// the region is added to the global grid once, z-contiguous.  An atom outside
// the region (never, while the rebuild criterion holds) adds directly.
constexpr int kSpreadTileCols = 144;    // columns per block: at most 12 x 12 (the host's choice)

template <class P>
__global__ __launch_bounds__(256) void pme_spread_tiled(const SysDev* __restrict__ sd,
                                                        const SysCount* __restrict__ sc,
                                                        const int4* __restrict__ tiles,
                                                        const int* __restrict__ col_count,
                                                        const int* __restrict__ col_slot,
                                                        const float* __restrict__ slot_zkey,
                                                        const typename P::real4* __restrict__ pos,
                                                        typename P::gacc* rgrid) {
    side_priority();
    using real = typename P::real;
    using acc = typename P::gacc;
    __shared__ acc g[kSpreadLdsPoints];
    __shared__ int c_first[kSpreadTileCols], c_cnt[kSpreadTileCols], c_a[kSpreadTileCols], c_e[kSpreadTileCols];
    __shared__ int c_pre[kSpreadTileCols + 1];
    const int4 t = tiles[blockIdx.x];          // (sys, bx, by, bz)
    const int s = t.x;
    const SysDev& d = sd[s];
    if (sc[s].n_slots == 0 || sc[s].overflow) return;     // overflowed build: the system is frozen
    const int Kx = d.grid[0], Ky = d.grid[1], Kz = d.grid[2];
    const int BX = d.sp_block[0], BY = d.sp_block[1], NZ = d.sp_block[2];
    const int cx0 = t.y * BX, cx1 = min(d.ncx, cx0 + BX);
    const int cy0 = t.z * BY, cy1 = min(d.ncy, cy0 + BY);
    const double zs = d.L[2] / NZ;
    const double z0 = t.w * zs, z1 = (t.w + 1 == NZ) ? d.L[2] + 1.0 : (t.w + 1) * zs;
    const double zlo = t.w * zs;
    const double m = d.sp_margin;
    // region origin and extent per axis (full axis when the block spans it)
    int lo[3], n[3];
    {
        const double a0[3] = {cx0 * d.colw[0], cy0 * d.colw[1], zlo};
        const double a1[3] = {cx1 * d.colw[0], cy1 * d.colw[1], (t.w + 1) * zs};
        const bool full[3] = {cx0 == 0 && cx1 == d.ncx, cy0 == 0 && cy1 == d.ncy, NZ == 1};
        const int K[3] = {Kx, Ky, Kz};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double h = d.L[k] / K[k];
            int l = (int)floor((a0[k] - m) / h);
            int u = (int)floor((a1[k] + m) / h) + 4;
            if (full[k] || u - l + 1 >= K[k]) { l = 0; u = K[k] - 1; }
            lo[k] = l;
            n[k] = u - l + 1;
        }
    }
    const bool fx = (n[0] == Kx && lo[0] == 0), fy = (n[1] == Ky && lo[1] == 0), fz = (n[2] == Kz && lo[2] == 0);
    const int pl = spread_plane(n[1], n[2]), nyz = n[1] * n[2];     // x-plane stride in LDS, points per plane
    const int npts = n[0] * pl;
    if (npts > kSpreadLdsPoints) return;       // host guarantees the fit
    for (int i = threadIdx.x; i < npts; i += blockDim.x) g[i] = acc(0);
    __syncthreads();
    // The block's columns (at most kSpreadTileCols) and, per column, the slots
    // whose build-time z lies in the slab: the slab bounds are found where the
    // column's sorted keys step across them, by all threads at once
    // (independent loads, no per-column search chain).
    const int ncy_t = cy1 - cy0, ncol = (cx1 - cx0) * ncy_t;
    for (int q = threadIdx.x; q < ncol; q += blockDim.x) {
        const int col = d.col_off + (cx0 + q / ncy_t) * d.ncy + cy0 + q % ncy_t;
        c_first[q] = col_slot[col];
        c_cnt[q] = col_count[col];
        c_a[q] = 0;
        c_e[q] = 0;
    }
    __syncthreads();
    {
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
        for (int q = wv; q < ncol; q += nw) {
            const int first = c_first[q], cnt = c_cnt[q];
            for (int i = lane; i < cnt; i += 64) {
                const double k1 = (double)slot_zkey[first + i];
                const double k0 = (i > 0) ? (double)slot_zkey[first + i - 1] : -1.0e300;
                if (k0 < z0 && z0 <= k1) c_a[q] = i;
                if (k0 < z1 && z1 <= k1) c_e[q] = i;
                if (i == cnt - 1) {
                    if (k1 < z0) c_a[q] = cnt;
                    if (k1 < z1) c_e[q] = cnt;
                }
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {     // running count of the columns' slab atoms
        int t = 0;
        for (int q = 0; q < ncol; ++q) {
            c_pre[q] = t;
            t += c_e[q] - c_a[q];
        }
        c_pre[ncol] = t;
    }
    __syncthreads();
    const int nat = c_pre[ncol];
    // 5 threads per atom (one per x plane), over all the block's columns
    for (int w = threadIdx.x; w < nat * 5; w += blockDim.x) {
        const int at = w / 5, ix = w - at * 5;
        int lo_q = 0, hi_q = ncol - 1;          // column q with c_pre[q] <= at < c_pre[q + 1]
        while (lo_q < hi_q) {
            const int mid = (lo_q + hi_q + 1) >> 1;
            if (c_pre[mid] <= at) lo_q = mid; else hi_q = mid - 1;
        }
        const int slot = c_first[lo_q] + c_a[lo_q] + (at - c_pre[lo_q]);
        const typename P::real4 p = pos[slot];
        int i0, i1, i2;
        real f0, f1, f2;
        grid_index<real>(p.x, d.invL[0], Kx, i0, f0);
        grid_index<real>(p.y, d.invL[1], Ky, i1, f1);
        grid_index<real>(p.z, d.invL[2], Kz, i2, f2);
        real tx[5], dtx[5], ty[5], dty[5], tz[5], dtz[5];
        bspline5<real>(f0, tx, dtx);
        bspline5<real>(f1, ty, dty);
        bspline5<real>(f2, tz, dtz);
        int u0 = fx ? i0 : i0 - lo[0];
        int u1 = fy ? i1 : i1 - lo[1];
        int u2 = fz ? i2 : i2 - lo[2];
        if (!fx) { if (u0 < 0) u0 += Kx; else if (u0 >= Kx) u0 -= Kx; }
        if (!fy) { if (u1 < 0) u1 += Ky; else if (u1 >= Ky) u1 -= Ky; }
        if (!fz) { if (u2 < 0) u2 += Kz; else if (u2 >= Kz) u2 -= Kz; }
        const bool inside = (fx || (u0 >= 0 && u0 + 4 < n[0])) && (fy || (u1 >= 0 && u1 + 4 < n[1])) &&
                            (fz || (u2 >= 0 && u2 + 4 < n[2]));
        const real qx = p.w * tx[ix];
        if (inside) {
            int x = u0 + ix;
            if (fx && x >= Kx) x -= Kx;
#pragma unroll
            for (int iy = 0; iy < 5; ++iy) {
                int y = u1 + iy;
                if (fy && y >= Ky) y -= Ky;
                const real qxy = qx * ty[iy];
#pragma unroll
                for (int iz = 0; iz < 5; ++iz) {
                    int z = u2 + iz;
                    if (fz && z >= Kz) z -= Kz;
                    grid_add_lds(&g[x * pl + y * n[2] + z], qxy * tz[iz]);
                }
            }
        } else {
            const int xi = (i0 + ix) % Kx;
            acc* gg = rgrid + d.rgrid_off + (long long)xi * Ky * Kz;
#pragma unroll
            for (int iy = 0; iy < 5; ++iy) {
                const int yi = (i1 + iy) % Ky;
                const real qxy = qx * ty[iy];
#pragma unroll
                for (int iz = 0; iz < 5; ++iz) grid_add(&gg[(long long)yi * Kz + (i2 + iz) % Kz], qxy * tz[iz]);
            }
        }
    }
    __syncthreads();
    // flush the region (z fastest: contiguous global addresses)
    const float inv_pl = 1.0f / (float)pl, inv_nz = 1.0f / (float)n[2];
    for (int i = threadIdx.x; i < npts; i += blockDim.x) {
        const acc v = g[i];
        if (v == acc(0)) continue;             // also the plane padding, never written
        const int x = qdiv(i, inv_pl), r = i - x * pl;     // i, pl <= kSpreadLdsPoints
        if (r >= nyz) continue;
        const int y = qdiv(r, inv_nz), z = r - y * n[2];
        int gx = lo[0] + x, gy = lo[1] + y, gz = lo[2] + z;
        gx = gx < 0 ? gx + Kx : (gx >= Kx ? gx - Kx : gx);
        gy = gy < 0 ? gy + Ky : (gy >= Ky ? gy - Ky : gy);
        gz = gz < 0 ? gz + Kz : (gz >= Kz ? gz - Kz : gz);
        grid_add_acc(&rgrid[d.rgrid_off + ((long long)gx * Ky + gy) * Kz + gz], v);
    }
}

// Force from the potential grid (NonbondedForce.cpp:407-461).  A block of
// 5 wavefronts covers 64 slots: wavefront ix takes stencil plane i0 + ix of
// every slot (25 grid loads per lane instead of 125, five times the loads in
// flight), and the five partial gradients are summed in LDS in plane order.
constexpr int kGatherSlots = 64;

template <class P>
__global__ __launch_bounds__(5 * kGatherSlots) void pme_gather_kernel(const SysDev* __restrict__ sd,
                                                                     const SysCount* __restrict__ sc,
                                                                     const int* __restrict__ slot_sys,
                                                                     const int* __restrict__ atom_of_slot,
                                                                     const typename P::real4* __restrict__ pos,
                                                                     const typename P::real* __restrict__ phi,
                                                                     typename P::facc* FX, typename P::facc* FY,
                                                                     typename P::facc* FZ, int n_slots_total) {
    side_priority();
    using real = typename P::real;
    __shared__ real part[3][5][kGatherSlots];
    const int ix = threadIdx.x / kGatherSlots, l = threadIdx.x - ix * kGatherSlots;
    const int slot = blockIdx.x * kGatherSlots + l;
    bool on = slot < n_slots_total;
    int s = 0;
    if (on) {
        s = slot_sys[slot];
        on = slot - sd[s].slot_off < sc[s].n_slots && !sc[s].overflow && atom_of_slot[slot] >= 0 &&
             sd[s].grid[0] != 0;
    }
    real fx = 0, fy = 0, fz = 0;
    typename P::real4 p;
    if (on) {
        const SysDev& d = sd[s];
        p = pos[slot];
        const int Kx = d.grid[0], Ky = d.grid[1], Kz = d.grid[2];
        int i0, i1, i2;
        real f0, f1, f2;
        grid_index<real>(p.x, d.invL[0], Kx, i0, f0);
        grid_index<real>(p.y, d.invL[1], Ky, i1, f1);
        grid_index<real>(p.z, d.invL[2], Kz, i2, f2);
        real tx[5], dtx[5], ty[5], dty[5], tz[5], dtz[5];
        bspline5<real>(f0, tx, dtx);
        bspline5<real>(f1, ty, dty);
        bspline5<real>(f2, tz, dtz);
        // stencil indices, periodic (i + t < 2K: one conditional subtraction, no division)
        int zi[5];
#pragma unroll
        for (int t = 0; t < 5; ++t) zi[t] = (i2 + t >= Kz) ? i2 + t - Kz : i2 + t;
        const int xi = (i0 + ix >= Kx) ? i0 + ix - Kx : i0 + ix;
        const real* g = phi + d.rgrid_off;
        real tix = tx[0], dtix = dtx[0];
#pragma unroll
        for (int k = 1; k < 5; ++k)
            if (ix == k) { tix = tx[k]; dtix = dtx[k]; }
#pragma unroll
        for (int iy = 0; iy < 5; ++iy) {
            const int yi = (i1 + iy >= Ky) ? i1 + iy - Ky : i1 + iy;
            const real* row = g + ((long long)xi * Ky + yi) * Kz;
            const real a = dtix * ty[iy], b = tix * dty[iy], c = tix * ty[iy];
#pragma unroll
            for (int iz = 0; iz < 5; ++iz) {
                const real v = row[zi[iz]];
                fx += a * tz[iz] * v;
                fy += b * tz[iz] * v;
                fz += c * dtz[iz] * v;
            }
        }
    }
    part[0][ix][l] = fx;
    part[1][ix][l] = fy;
    part[2][ix][l] = fz;
    __syncthreads();
    if (ix == 0 && on) {
        const SysDev& d = sd[s];
        real gx = 0, gy = 0, gz = 0;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            gx += part[0][k][l];
            gy += part[1][k][l];
            gz += part[2][k][l];
        }
        const real q = p.w;
        facc_add(&FX[slot], -q * (gx * (real)d.grid[0] * (real)d.invL[0]));
        facc_add(&FY[slot], -q * (gy * (real)d.grid[1] * (real)d.invL[1]));
        facc_add(&FZ[slot], -q * (gz * (real)d.grid[2] * (real)d.invL[2]));
    }
}

// ------------------------------------------------------------------ FFT

template <typename R2>
__device__ __forceinline__ R2 cmul(R2 a, R2 b) {
    R2 r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}

// Small DFT of radix r in registers; sign -1 forward, +1 backward.
template <typename real, typename R2>
__device__ __forceinline__ void dft_small(R2* v, int r, int sign) {
    if (r == 2) {
        R2 a = v[0], b = v[1];
        v[0].x = a.x + b.x; v[0].y = a.y + b.y;
        v[1].x = a.x - b.x; v[1].y = a.y - b.y;
    } else if (r == 4) {
        R2 a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3];
        R2 s02 = {a0.x + a2.x, a0.y + a2.y}, d02 = {a0.x - a2.x, a0.y - a2.y};
        R2 s13 = {a1.x + a3.x, a1.y + a3.y}, d13 = {a1.x - a3.x, a1.y - a3.y};
        // W = sign * i ; W * (x + iy) = sign * (-y + ix)
        R2 wd = {(real)sign * -d13.y, (real)sign * d13.x};
        v[0].x = s02.x + s13.x; v[0].y = s02.y + s13.y;
        v[2].x = s02.x - s13.x; v[2].y = s02.y - s13.y;
        v[1].x = d02.x + wd.x;  v[1].y = d02.y + wd.y;
        v[3].x = d02.x - wd.x;  v[3].y = d02.y - wd.y;
    } else if (r == 3) {
        const real c1 = (real)-0.5, s1 = (real)sign * (real)0.86602540378443864676;
        R2 a0 = v[0], a1 = v[1], a2 = v[2];
        R2 sm = {a1.x + a2.x, a1.y + a2.y}, df = {a1.x - a2.x, a1.y - a2.y};
        R2 t = {a0.x + c1 * sm.x, a0.y + c1 * sm.y};
        // i * s1 * df = (-s1 df.y, s1 df.x)
        v[0].x = a0.x + sm.x; v[0].y = a0.y + sm.y;
        v[1].x = t.x - s1 * df.y; v[1].y = t.y + s1 * df.x;
        v[2].x = t.x + s1 * df.y; v[2].y = t.y - s1 * df.x;
    } else {  // r == 5
        const real c1 = (real)0.30901699437494742410, c2 = (real)-0.80901699437494742410;
        const real s1 = (real)sign * (real)0.95105651629515357212, s2 = (real)sign * (real)0.58778525229247312917;
        R2 a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3], a4 = v[4];
        R2 b1 = {a1.x + a4.x, a1.y + a4.y}, b4 = {a1.x - a4.x, a1.y - a4.y};
        R2 b2 = {a2.x + a3.x, a2.y + a3.y}, b3 = {a2.x - a3.x, a2.y - a3.y};
        R2 t1 = {a0.x + c1 * b1.x + c2 * b2.x, a0.y + c1 * b1.y + c2 * b2.y};
        R2 t2 = {a0.x + c2 * b1.x + c1 * b2.x, a0.y + c2 * b1.y + c1 * b2.y};
        // u1 = s1 b4 + s2 b3, u2 = s2 b4 - s1 b3 ; X1 = t1 + i u1, X4 = t1 - i u1, X2 = t2 + i u2, X3 = t2 - i u2
        R2 u1 = {s1 * b4.x + s2 * b3.x, s1 * b4.y + s2 * b3.y};
        R2 u2 = {s2 * b4.x - s1 * b3.x, s2 * b4.y - s1 * b3.y};
        v[0].x = a0.x + b1.x + b2.x; v[0].y = a0.y + b1.y + b2.y;
        v[1].x = t1.x - u1.y; v[1].y = t1.y + u1.x;
        v[4].x = t1.x + u1.y; v[4].y = t1.y - u1.x;
        v[2].x = t2.x - u2.y; v[2].y = t2.y + u2.x;
        v[3].x = t2.x + u2.y; v[3].y = t2.y - u2.x;
    }
}

__device__ __forceinline__ int next_radix(int m) {
    if (m % 4 == 0) return 4;
    if (m % 2 == 0) return 2;
    if (m % 3 == 0) return 3;
    return 5;
}


// Copies the N twiddles of an axis into LDS (all threads of the block).
template <typename R2>
__device__ __forceinline__ void load_twiddles(R2* dst, const R2* __restrict__ src, int N) {
    for (int i = threadIdx.x; i < N; i += blockDim.x) dst[i] = src[i];
}

// FFT of `lines` lines of length N in LDS, point m of line l at a[l * ls + m * es];
// `b` is scratch of the same layout; `tw` holds exp(-2 pi i m / N), m < N, in
// LDS.  Stockham autosort, radix 4, 2, 3, 5; unnormalised; sign -1 forward, +1
// backward.  Returns the buffer holding the result.  All threads of the block
// call it; every stage ends with a barrier.
template <typename real, typename R2>
__device__ R2* fft_lines_lds(R2* a, R2* b, int N, int lines, int ls, int es, const R2* tw, int sign) {
    int Ns = 1;
    int rem = N;
    R2* src = a;
    R2* dst = b;
    while (rem > 1) {
        const int r = next_radix(rem);
        const int nb = N / r;
        const int step = N / (Ns * r);
        const float inv_nb = 1.0f / (float)nb, inv_ns = 1.0f / (float)Ns;
        const int total = lines * nb;
        for (int w = threadIdx.x; w < total; w += blockDim.x) {
            const int line = qdiv(w, inv_nb), j = w - line * nb;
            const int jq = qdiv(j, inv_ns), k = j - jq * Ns;
            const R2* sp = src + line * ls;
            R2* dp = dst + line * ls;
            R2 v[5];
#pragma unroll
            for (int q = 0; q < 5; ++q)
                if (q < r) v[q] = sp[(j + q * nb) * es];
            if (k != 0) {
#pragma unroll
                for (int q = 1; q < 5; ++q) {
                    if (q < r) {
                        R2 t = tw[k * q * step];     // k q step < N
                        if (sign > 0) t.y = -t.y;
                        v[q] = cmul(v[q], t);
                    }
                }
            }
            dft_small<real, R2>(v, r, sign);
            const int base = jq * Ns * r + k;
#pragma unroll
            for (int q = 0; q < 5; ++q)
                if (q < r) dp[(base + q * Ns) * es] = v[q];
        }
        __syncthreads();
        R2* t = src;
        src = dst;
        dst = t;
        Ns *= r;
        rem /= r;
    }
    return src;
}

__device__ __forceinline__ float fixed_to_real(long long v, float) { return (float)((double)v * (1.0 / PrecMixed::grid_scale)); }
__device__ __forceinline__ double fixed_to_real(long long v, double) { return (double)v * (1.0 / PrecFP64::grid_scale); }

// Z pass forward: R2C.  item = (sys, first line, n lines)
template <class P>
__global__ __launch_bounds__(256) void fft_z_fwd(const SysDev* __restrict__ sd, const int4* __restrict__ items,
                                                 typename P::gacc* __restrict__ rgrid,
                                                 typename P::real2* cgrid, const typename P::real2* __restrict__ tw) {
    side_priority();
    using real = typename P::real;
    using R2 = typename P::real2;
    extern __shared__ unsigned char smem_raw[];
    R2* smem = reinterpret_cast<R2*>(smem_raw);
    const int4 it = items[blockIdx.x];
    const SysDev& d = sd[it.x];
    const int Kz = d.grid[2], Kzh = Kz / 2 + 1;
    const int line0 = it.y, nl = it.z;
    R2* A = smem;
    R2* B = smem + nl * Kz;
    R2* TW = B + nl * Kz;
    load_twiddles(TW, tw + d.tw_off[2], Kz);
    // the spread grid is cleared as it is read: the next spread adds into it
    typename P::gacc* g = rgrid + d.rgrid_off + (long long)line0 * Kz;
    for (int i = threadIdx.x; i < nl * Kz; i += blockDim.x) {
        R2 v;
        v.x = fixed_to_real(g[i], (real)0);
        v.y = 0;
        g[i] = typename P::gacc(0);
        A[i] = v;
    }
    __syncthreads();
    R2* res = fft_lines_lds<real, R2>(A, B, Kz, nl, Kz, 1, TW, -1);
    R2* c = cgrid + d.cgrid_off + (long long)line0 * Kzh;
    const float inv_kzh = 1.0f / (float)Kzh;
    for (int i = threadIdx.x; i < nl * Kzh; i += blockDim.x) {
        const int t = qdiv(i, inv_kzh), kz = i - t * Kzh;
        c[i] = res[t * Kz + kz];
    }
}

// Z pass inverse: C2R into the potential grid.
template <class P>
__global__ __launch_bounds__(256) void fft_z_inv(const SysDev* __restrict__ sd, const int4* __restrict__ items,
                                                 const typename P::real2* __restrict__ cgrid, typename P::real* phi,
                                                 const typename P::real2* __restrict__ tw) {
    side_priority();
    using real = typename P::real;
    using R2 = typename P::real2;
    extern __shared__ unsigned char smem_raw[];
    R2* smem = reinterpret_cast<R2*>(smem_raw);
    const int4 it = items[blockIdx.x];
    const SysDev& d = sd[it.x];
    const int Kz = d.grid[2], Kzh = Kz / 2 + 1;
    const int line0 = it.y, nl = it.z;
    R2* A = smem;
    R2* B = smem + nl * Kz;
    R2* TW = B + nl * Kz;
    load_twiddles(TW, tw + d.tw_off[2], Kz);
    const R2* c = cgrid + d.cgrid_off + (long long)line0 * Kzh;
    const float inv_kz = 1.0f / (float)Kz;
    for (int i = threadIdx.x; i < nl * Kz; i += blockDim.x) {
        const int t = qdiv(i, inv_kz), z = i - t * Kz;
        R2 v;
        if (z < Kzh) v = c[t * Kzh + z];
        else { v = c[t * Kzh + (Kz - z)]; v.y = -v.y; }
        A[i] = v;
    }
    __syncthreads();
    R2* res = fft_lines_lds<real, R2>(A, B, Kz, nl, Kz, 1, TW, +1);
    real* o = phi + d.rgrid_off + (long long)line0 * Kz;
    for (int i = threadIdx.x; i < nl * Kz; i += blockDim.x) o[i] = res[i].x;
}

// Y pass: tile (sys, x, kz0) of TZ consecutive kz columns, lines along y.
template <class P>
__global__ __launch_bounds__(256) void fft_y(const SysDev* __restrict__ sd, const int4* __restrict__ items,
                                             typename P::real2* cgrid, const typename P::real2* __restrict__ tw,
                                             int TZ, int sign) {
    side_priority();
    using real = typename P::real;
    using R2 = typename P::real2;
    extern __shared__ unsigned char smem_raw[];
    R2* smem = reinterpret_cast<R2*>(smem_raw);
    const int4 it = items[blockIdx.x];
    const SysDev& d = sd[it.x];
    const int Ky = d.grid[1], Kzh = d.grid[2] / 2 + 1;
    const int x = it.y, kz0 = it.z;
    const int nt = min(TZ, Kzh - kz0);
    R2* A = smem;
    R2* B = smem + TZ * Ky;
    R2* TW = B + TZ * Ky;
    load_twiddles(TW, tw + d.tw_off[1], Ky);
    R2* c = cgrid + d.cgrid_off + (long long)x * Ky * Kzh + kz0;
    const float inv_nt = 1.0f / (float)nt;
    for (int i = threadIdx.x; i < nt * Ky; i += blockDim.x) {
        const int y = qdiv(i, inv_nt), t = i - y * nt;
        A[t * Ky + y] = c[(long long)y * Kzh + t];
    }
    __syncthreads();
    R2* res = fft_lines_lds<real, R2>(A, B, Ky, nt, Ky, 1, TW, sign);
    for (int i = threadIdx.x; i < nt * Ky; i += blockDim.x) {
        const int y = qdiv(i, inv_nt), t = i - y * nt;
        c[(long long)y * Kzh + t] = res[t * Ky + y];
    }
}

// X pass: forward FFT along x, convolution with eta (and energy), inverse FFT.
template <class P, bool ENERGY>
__global__ __launch_bounds__(256) void fft_x_conv(const SysDev* __restrict__ sd, const int4* __restrict__ items,
                                                  typename P::real2* cgrid, const typename P::real* __restrict__ eta,
                                                  const typename P::real2* __restrict__ tw, int TZ,
                                                  long long* energy) {
    side_priority();
    using real = typename P::real;
    using R2 = typename P::real2;
    extern __shared__ unsigned char smem_raw[];
    R2* smem = reinterpret_cast<R2*>(smem_raw);
    const int4 it = items[blockIdx.x];
    const int s = it.x;
    const SysDev& d = sd[s];
    const int Kx = d.grid[0], Ky = d.grid[1], Kz = d.grid[2], Kzh = Kz / 2 + 1;
    const int y = it.y, kz0 = it.z;
    const int nt = min(TZ, Kzh - kz0);
    R2* A = smem;
    R2* B = smem + TZ * Kx;
    const long long sx = (long long)Ky * Kzh;
    R2* c = cgrid + d.cgrid_off + (long long)y * Kzh + kz0;
    const real* et = eta + d.cgrid_off + (long long)y * Kzh + kz0;
    R2* TW = B + TZ * Kx;
    load_twiddles(TW, tw + d.tw_off[0], Kx);
    const float inv_nt = 1.0f / (float)nt;
    for (int i = threadIdx.x; i < nt * Kx; i += blockDim.x) {
        const int x = qdiv(i, inv_nt), t = i - x * nt;
        A[t * Kx + x] = c[x * sx + t];
    }
    __syncthreads();
    R2* res = fft_lines_lds<real, R2>(A, B, Kx, nt, Kx, 1, TW, -1);
    long long e = 0;   // fixed point per element: independent of the tile shape
    for (int i = threadIdx.x; i < nt * Kx; i += blockDim.x) {
        const int x = qdiv(i, inv_nt), t = i - x * nt;
        const real w = eta ? et[x * sx + t] : (real)0;
        R2 v = res[t * Kx + x];
        if constexpr (ENERGY) {
            const int kz = kz0 + t;
            const double wt = (kz == 0 || (2 * kz == Kz)) ? 0.5 : 1.0;
            e += energy_fixed(wt * (double)w * ((double)v.x * v.x + (double)v.y * v.y));
        }
        v.x *= w;
        v.y *= w;
        res[t * Kx + x] = v;
    }
    __syncthreads();
    R2* other = (res == A) ? B : A;
    R2* res2 = fft_lines_lds<real, R2>(res, other, Kx, nt, Kx, 1, TW, +1);
    for (int i = threadIdx.x; i < nt * Kx; i += blockDim.x) {
        const int x = qdiv(i, inv_nt), t = i - x * nt;
        c[x * sx + t] = res2[t * Kx + x];
    }
    if constexpr (ENERGY) block_energy_add(&energy[s * E_NTERMS + E_RECIP], e);
}

// Plain X pass (no convolution) for the standalone transform.
template <class P>
__global__ __launch_bounds__(256) void fft_x_plain(const SysDev* __restrict__ sd, const int4* __restrict__ items,
                                                   typename P::real2* cgrid, const typename P::real2* __restrict__ tw,
                                                   int TZ, int sign) {
    side_priority();
    using real = typename P::real;
    using R2 = typename P::real2;
    extern __shared__ unsigned char smem_raw[];
    R2* smem = reinterpret_cast<R2*>(smem_raw);
    const int4 it = items[blockIdx.x];
    const SysDev& d = sd[it.x];
    const int Kx = d.grid[0], Ky = d.grid[1], Kzh = d.grid[2] / 2 + 1;
    const int y = it.y, kz0 = it.z;
    const int nt = min(TZ, Kzh - kz0);
    R2* A = smem;
    R2* B = smem + TZ * Kx;
    const long long sx = (long long)Ky * Kzh;
    R2* c = cgrid + d.cgrid_off + (long long)y * Kzh + kz0;
    R2* TW = B + TZ * Kx;
    load_twiddles(TW, tw + d.tw_off[0], Kx);
    const float inv_nt = 1.0f / (float)nt;
    for (int i = threadIdx.x; i < nt * Kx; i += blockDim.x) {
        const int x = qdiv(i, inv_nt), t = i - x * nt;
        A[t * Kx + x] = c[x * sx + t];
    }
    __syncthreads();
    R2* res = fft_lines_lds<real, R2>(A, B, Kx, nt, Kx, 1, TW, sign);
    for (int i = threadIdx.x; i < nt * Kx; i += blockDim.x) {
        const int x = qdiv(i, inv_nt), t = i - x * nt;
        c[x * sx + t] = res[t * Kx + x];
    }
}

// ---------------------------------------------------------- plane FFT passes
//
// Whole x-planes in LDS (gfx950 gives a workgroup up to 160 KiB): the forward
// pass transforms along z and then y in one kernel, the x pass transforms
// along x, multiplies by the influence function and transforms back, and the
// inverse pass runs y and then z.  Three launches per step instead of five.
// The real-to-complex z transform handles two real lines (y = 2p, 2p + 1) as
// one complex line a + ib and separates them afterwards:
//   A_k = (Z_k + conj Z_{N-k}) / 2,  B_k = (Z_k - conj Z_{N-k}) / 2i.
// The complex-to-real inverse packs them the same way: Z_k = A_k + i B_k over
// the Hermitian extension, then a = Re, b = Im of the inverse transform.

// Forward: plane x of the spread grid -> plane x of the half spectrum (z then y).
// item = (sys, x)
template <class P>
__global__ __launch_bounds__(kPlaneThreads) void fft_plane_fwd(const SysDev* __restrict__ sd,
                                                              const int4* __restrict__ items,
                                                              typename P::gacc* __restrict__ rgrid,
                                                              typename P::real2* cgrid,
                                                              const typename P::real2* __restrict__ tw) {
    side_priority();
    using real = typename P::real;
    using R2 = typename P::real2;
    extern __shared__ unsigned char smem_raw[];
    R2* smem = reinterpret_cast<R2*>(smem_raw);
    const int4 it = items[blockIdx.x];
    const SysDev& d = sd[it.x];
    const int x = it.y;
    const int Ky = d.grid[1], Kz = d.grid[2], Kzh = Kz / 2 + 1, np = (Ky + 1) >> 1;
    const int cap = max(np * Kz, Ky * Kzh);
    R2* TWy = smem;
    R2* TWz = TWy + Ky;
    R2* A = TWz + Kz;
    R2* Bf = A + cap;
    load_twiddles(TWy, tw + d.tw_off[1], Ky);
    load_twiddles(TWz, tw + d.tw_off[2], Kz);
    // the spread grid is cleared as it is read: the next spread adds into it
    typename P::gacc* g = rgrid + d.rgrid_off + (long long)x * Ky * Kz;
    const float inv_kz = 1.0f / (float)Kz;
    for (int i = threadIdx.x; i < np * Kz; i += blockDim.x) {
        const int p = qdiv(i, inv_kz), z = i - p * Kz;
        R2 v;
        v.x = fixed_to_real(g[(2 * p) * Kz + z], (real)0);
        g[(2 * p) * Kz + z] = typename P::gacc(0);
        v.y = (real)0;
        if (2 * p + 1 < Ky) {
            v.y = fixed_to_real(g[(2 * p + 1) * Kz + z], (real)0);
            g[(2 * p + 1) * Kz + z] = typename P::gacc(0);
        }
        A[i] = v;
    }
    __syncthreads();
    R2* Z = fft_lines_lds<real, R2>(A, Bf, Kz, np, Kz, 1, TWz, -1);
    R2* Q = (Z == A) ? Bf : A;
    const float inv_kzh = 1.0f / (float)Kzh;
    for (int i = threadIdx.x; i < np * Kzh; i += blockDim.x) {
        const int p = qdiv(i, inv_kzh), k = i - p * Kzh;
        const R2 zk = Z[p * Kz + k], zn = Z[p * Kz + (k == 0 ? 0 : Kz - k)];
        R2 a, b;
        a.x = (real)0.5 * (zk.x + zn.x);
        a.y = (real)0.5 * (zk.y - zn.y);
        b.x = (real)0.5 * (zk.y + zn.y);
        b.y = (real)0.5 * (zn.x - zk.x);
        Q[(2 * p) * Kzh + k] = a;
        if (2 * p + 1 < Ky) Q[(2 * p + 1) * Kzh + k] = b;
    }
    __syncthreads();
    R2* Y = fft_lines_lds<real, R2>(Q, Z, Ky, Kzh, 1, Kzh, TWy, -1);
    R2* c = cgrid + d.cgrid_off + (long long)x * Ky * Kzh;       // the plane is contiguous
    for (int i = threadIdx.x; i < Ky * Kzh; i += blockDim.x) c[i] = Y[i];
}

// Inverse: plane x of the half spectrum -> plane x of the potential (y then z).
template <class P>
__global__ __launch_bounds__(kPlaneThreads) void fft_plane_inv(const SysDev* __restrict__ sd,
                                                              const int4* __restrict__ items,
                                                              const typename P::real2* __restrict__ cgrid,
                                                              typename P::real* phi,
                                                              const typename P::real2* __restrict__ tw) {
    side_priority();
    using real = typename P::real;
    using R2 = typename P::real2;
    extern __shared__ unsigned char smem_raw[];
    R2* smem = reinterpret_cast<R2*>(smem_raw);
    const int4 it = items[blockIdx.x];
    const SysDev& d = sd[it.x];
    const int x = it.y;
    const int Ky = d.grid[1], Kz = d.grid[2], Kzh = Kz / 2 + 1, np = (Ky + 1) >> 1;
    const int cap = max(np * Kz, Ky * Kzh);
    R2* TWy = smem;
    R2* TWz = TWy + Ky;
    R2* A = TWz + Kz;
    R2* Bf = A + cap;
    load_twiddles(TWy, tw + d.tw_off[1], Ky);
    load_twiddles(TWz, tw + d.tw_off[2], Kz);
    const R2* c = cgrid + d.cgrid_off + (long long)x * Ky * Kzh;
    for (int i = threadIdx.x; i < Ky * Kzh; i += blockDim.x) A[i] = c[i];
    __syncthreads();
    R2* Y = fft_lines_lds<real, R2>(A, Bf, Ky, Kzh, 1, Kzh, TWy, +1);
    R2* W = (Y == A) ? Bf : A;
    const float inv_kz = 1.0f / (float)Kz;
    for (int i = threadIdx.x; i < np * Kz; i += blockDim.x) {
        const int p = qdiv(i, inv_kz), k = i - p * Kz;
        const bool hi = k >= Kzh;                 // upper half from the Hermitian symmetry
        const int ks = hi ? Kz - k : k;
        R2 a = Y[(2 * p) * Kzh + ks];
        R2 b;
        if (2 * p + 1 < Ky) b = Y[(2 * p + 1) * Kzh + ks];
        else { b.x = 0; b.y = 0; }
        if (ks == 0 || 2 * ks == Kz) { a.y = 0; b.y = 0; }      // DC and Nyquist terms are real
        if (hi) { a.y = -a.y; b.y = -b.y; }
        R2 zv;
        zv.x = a.x - b.y;
        zv.y = a.y + b.x;
        W[i] = zv;
    }
    __syncthreads();
    R2* Z = fft_lines_lds<real, R2>(W, Y, Kz, np, Kz, 1, TWz, +1);
    real* o = phi + d.rgrid_off + (long long)x * Ky * Kz;
    for (int i = threadIdx.x; i < np * Kz; i += blockDim.x) {
        const int p = qdiv(i, inv_kz), z = i - p * Kz;
        o[(2 * p) * Kz + z] = Z[i].x;
        if (2 * p + 1 < Ky) o[(2 * p + 1) * Kz + z] = Z[i].y;
    }
}

// X pass over the lines (y, k0 .. k0 + nk - 1): MODE 0 forward, 1 inverse,
// 2 forward / influence function / inverse, 3 the same with the energy.
// item = (sys, y, k0, nk)
template <class P, int MODE>
__global__ __launch_bounds__(kPlaneThreads) void fft_x_block(const SysDev* __restrict__ sd,
                                                            const int4* __restrict__ items,
                                                            typename P::real2* cgrid,
                                                            const typename P::real* __restrict__ eta,
                                                            const typename P::real2* __restrict__ tw,
                                                            long long* energy) {
    side_priority();
    using real = typename P::real;
    using R2 = typename P::real2;
    extern __shared__ unsigned char smem_raw[];
    R2* smem = reinterpret_cast<R2*>(smem_raw);
    const int4 it = items[blockIdx.x];
    const int s = it.x, y = it.y, k0 = it.z, nk = it.w;
    const SysDev& d = sd[s];
    const int Kx = d.grid[0], Ky = d.grid[1], Kz = d.grid[2], Kzh = Kz / 2 + 1;
    R2* TW = smem;
    R2* A = TW + Kx;
    R2* Bf = A + nk * Kx;
    load_twiddles(TW, tw + d.tw_off[0], Kx);
    const long long sx = (long long)Ky * Kzh;
    R2* c = cgrid + d.cgrid_off + (long long)y * Kzh + k0;
    const float inv_nk = 1.0f / (float)nk;
    // A[x * nk + t]: one row of nk wave numbers per x
    for (int i = threadIdx.x; i < Kx * nk; i += blockDim.x) {
        const int x = qdiv(i, inv_nk), t = i - x * nk;
        A[i] = c[x * sx + t];
    }
    __syncthreads();
    R2* R = fft_lines_lds<real, R2>(A, Bf, Kx, nk, 1, nk, TW, MODE == 1 ? +1 : -1);
    if constexpr (MODE >= 2) {
        const real* et = eta + d.cgrid_off + (long long)y * Kzh + k0;
        long long e = 0;   // fixed point per element: independent of the tiling
        for (int i = threadIdx.x; i < Kx * nk; i += blockDim.x) {
            const int x = qdiv(i, inv_nk), t = i - x * nk;
            const real w = et[x * sx + t];
            R2 v = R[i];
            if constexpr (MODE == 3) {
                const int kz = k0 + t;
                const double wt = (kz == 0 || (2 * kz == Kz)) ? 0.5 : 1.0;
                e += energy_fixed(wt * (double)w * ((double)v.x * v.x + (double)v.y * v.y));
            }
            v.x *= w;
            v.y *= w;
            R[i] = v;
        }
        __syncthreads();
        R = fft_lines_lds<real, R2>(R, R == A ? Bf : A, Kx, nk, 1, nk, TW, +1);
        if constexpr (MODE == 3) block_energy_add(&energy[s * E_NTERMS + E_RECIP], e);
    }
    for (int i = threadIdx.x; i < Kx * nk; i += blockDim.x) {
        const int x = qdiv(i, inv_nk), t = i - x * nk;
        c[x * sx + t] = R[i];
    }
}

template <class P>
void launch_pme_spread(Batch<P>& B) {
    if (!B.pme_any) return;
    const int T = 256;
    // the grid starts cleared (allocation) and the forward transform clears it
    // again as it reads it, so every spread adds into zeros
    if (B.n_spread_tiles > 0)
        pme_spread_tiled<P><<<B.n_spread_tiles, 256, 0, B.stream>>>(B.d_sys, B.d_cnt, B.spread_tiles, B.col_count,
                                                                    B.col_slot, B.slot_zkey, B.pos, B.rgrid);
    if (B.n_spread_direct_slots > 0) {
        // systems whose region does not fit the LDS tile: per-atom global atomics
        const long long nthr = (long long)B.n_slots_cap * 5;
        pme_spread_kernel<P><<<(unsigned)((nthr + T - 1) / T), T, 0, B.stream>>>(B.d_sys, B.d_cnt, B.slot_sys,
                                                                                 B.atom_of_slot, B.pos, B.rgrid,
                                                                                 B.n_slots_cap);
    }
}

template <class P>
static size_t fft_smem(const Batch<P>& B, int lines, int N) {
    return (size_t)(2 * lines + 1) * N * sizeof(typename P::real2);   // two line buffers + twiddles
}

// NEXUSMD_FFT_PLANE=0 selects the five line passes even where planes fit.
bool fft_plane_enabled() {
    const char* e = std::getenv("NEXUSMD_FFT_PLANE");
    return !(e && std::string(e) == "0");
}

// Raises the dynamic-LDS limit of the plane kernels to what the batch needs.
template <class P>
hipError_t prepare_fft(Batch<P>& B) {
    if (!B.fft_plane) return hipSuccess;
    const int pl = (int)B.fft_plane_lds, xl = (int)B.fft_x_lds;
    const hipFuncAttribute a = hipFuncAttributeMaxDynamicSharedMemorySize;
    hipError_t e = hipFuncSetAttribute((const void*)fft_plane_fwd<P>, a, pl);
    if (e == hipSuccess) e = hipFuncSetAttribute((const void*)fft_plane_inv<P>, a, pl);
    if (e == hipSuccess) e = hipFuncSetAttribute((const void*)fft_x_block<P, 2>, a, xl);
    if (e == hipSuccess) e = hipFuncSetAttribute((const void*)fft_x_block<P, 3>, a, xl);
    return e;
}

template <class P>
void launch_fft_forward(Batch<P>& B) {
    if (!B.pme_any) return;
    if (B.fft_plane) {
        fft_plane_fwd<P><<<B.n_fftz, kPlaneThreads, B.fft_plane_lds, B.stream>>>(B.d_sys, B.fftz_items, B.rgrid,
                                                                               B.cgrid, B.twiddle);
        return;
    }
    fft_z_fwd<P><<<B.n_fftz, 256, fft_smem(B, B.fft_lines, B.fft_maxK), B.stream>>>(B.d_sys, B.fftz_items, B.rgrid,
                                                                                   B.cgrid, B.twiddle);
    fft_y<P><<<B.n_ffty, 256, fft_smem(B, B.fft_tile, B.fft_maxK), B.stream>>>(B.d_sys, B.ffty_items, B.cgrid,
                                                                              B.twiddle, B.fft_tile, -1);
}

template <class P>
void launch_fft_convolve(Batch<P>& B, bool energy) {
    if (!B.pme_any) return;
    if (B.fft_plane) {
        if (energy)
            fft_x_block<P, 3><<<B.n_fftx, kXPassThreads, B.fft_x_lds, B.stream>>>(B.d_sys, B.fftx_items, B.cgrid,
                                                                                B.eta, B.twiddle, B.energy);
        else
            fft_x_block<P, 2><<<B.n_fftx, kXPassThreads, B.fft_x_lds, B.stream>>>(B.d_sys, B.fftx_items, B.cgrid,
                                                                                B.eta, B.twiddle, B.energy);
        return;
    }
    if (energy)
        fft_x_conv<P, true><<<B.n_fftx, 256, fft_smem(B, B.fft_tile, B.fft_maxK), B.stream>>>(
            B.d_sys, B.fftx_items, B.cgrid, B.eta, B.twiddle, B.fft_tile, B.energy);
    else
        fft_x_conv<P, false><<<B.n_fftx, 256, fft_smem(B, B.fft_tile, B.fft_maxK), B.stream>>>(
            B.d_sys, B.fftx_items, B.cgrid, B.eta, B.twiddle, B.fft_tile, B.energy);
}

template <class P>
void launch_fft_inverse(Batch<P>& B) {
    if (!B.pme_any) return;
    if (B.fft_plane) {
        fft_plane_inv<P><<<B.n_fftz, kPlaneThreads, B.fft_plane_lds, B.stream>>>(B.d_sys, B.fftz_items, B.cgrid,
                                                                               B.phi, B.twiddle);
        return;
    }
    fft_y<P><<<B.n_ffty, 256, fft_smem(B, B.fft_tile, B.fft_maxK), B.stream>>>(B.d_sys, B.ffty_items, B.cgrid,
                                                                              B.twiddle, B.fft_tile, +1);
    fft_z_inv<P><<<B.n_fftz, 256, fft_smem(B, B.fft_lines, B.fft_maxK), B.stream>>>(B.d_sys, B.fftz_items, B.cgrid,
                                                                                   B.phi, B.twiddle);
}

template <class P>
void launch_pme_gather(Batch<P>& B) {
    if (!B.pme_any) return;
    pme_gather_kernel<P><<<(unsigned)((B.n_slots_cap + kGatherSlots - 1) / kGatherSlots), 5 * kGatherSlots, 0,
                           B.stream>>>(B.d_sys, B.d_cnt, B.slot_sys, B.atom_of_slot, B.pos, B.phi, B.FSx, B.FSy,
                                       B.FSz, B.n_slots_cap);
}

template void launch_pme_spread<PrecFP64>(Batch<PrecFP64>&);
template void launch_pme_spread<PrecMixed>(Batch<PrecMixed>&);
template hipError_t prepare_fft<PrecFP64>(Batch<PrecFP64>&);
template hipError_t prepare_fft<PrecMixed>(Batch<PrecMixed>&);
template void launch_fft_forward<PrecFP64>(Batch<PrecFP64>&);
template void launch_fft_forward<PrecMixed>(Batch<PrecMixed>&);
template void launch_fft_convolve<PrecFP64>(Batch<PrecFP64>&, bool);
template void launch_fft_convolve<PrecMixed>(Batch<PrecMixed>&, bool);
template void launch_fft_inverse<PrecFP64>(Batch<PrecFP64>&);
template void launch_fft_inverse<PrecMixed>(Batch<PrecMixed>&);
template void launch_pme_gather<PrecFP64>(Batch<PrecFP64>&);
template void launch_pme_gather<PrecMixed>(Batch<PrecMixed>&);

// Standalone 3D transform of one real grid through the same pass kernels.
template <class P>
void fft3d_roundtrip(int nx, int ny, int nz, const double* in, double* out, bool forward, hipStream_t st) {
    using R2 = typename P::real2;
    using real = typename P::real;
    const int nzh = nz / 2 + 1;
    SysDev d{};
    d.grid[0] = nx; d.grid[1] = ny; d.grid[2] = nz;
    d.rgrid_off = 0; d.cgrid_off = 0;
    d.tw_off[0] = 0; d.tw_off[1] = nx; d.tw_off[2] = nx + ny;
    std::vector<R2> tw(nx + ny + nz);
    const int K[3] = {nx, ny, nz};
    for (int a = 0, o = 0; a < 3; o += K[a], ++a)
        for (int t = 0; t < K[a]; ++t) {
            const double ang = -2.0 * kPi * (double)t / (double)K[a];
            tw[o + t].x = (real)cos(ang);
            tw[o + t].y = (real)sin(ang);
        }
    const int maxK = std::max(nx, std::max(ny, nz));
    const int T = std::max(1, std::min(8, (int)((61440 / (int)sizeof(R2) - maxK) / (2 * maxK))));
    std::vector<int4> zi, yi, xi;
    for (int l = 0; l < nx * ny; l += T) zi.push_back(make_int4(0, l, std::min(T, nx * ny - l), 0));
    for (int x = 0; x < nx; ++x)
        for (int k = 0; k < nzh; k += T) yi.push_back(make_int4(0, x, k, 0));
    for (int y = 0; y < ny; ++y)
        for (int k = 0; k < nzh; k += T) xi.push_back(make_int4(0, y, k, 0));
    SysDev* dd; R2* dtw; int4 *dz, *dy, *dx; typename P::gacc* rg; R2* cg; real* ph;
    const size_t G = (size_t)nx * ny * nz, Gh = (size_t)nx * ny * nzh;
    hipMalloc(&dd, sizeof(SysDev));
    hipMalloc(&dtw, tw.size() * sizeof(R2));
    hipMalloc(&dz, zi.size() * sizeof(int4));
    hipMalloc(&dy, yi.size() * sizeof(int4));
    hipMalloc(&dx, xi.size() * sizeof(int4));
    hipMalloc(&rg, G * sizeof(typename P::gacc));
    hipMalloc(&cg, Gh * sizeof(R2));
    hipMalloc(&ph, G * sizeof(real));
    hipMemcpyAsync(dd, &d, sizeof(SysDev), hipMemcpyHostToDevice, st);
    hipMemcpyAsync(dtw, tw.data(), tw.size() * sizeof(R2), hipMemcpyHostToDevice, st);
    hipMemcpyAsync(dz, zi.data(), zi.size() * sizeof(int4), hipMemcpyHostToDevice, st);
    hipMemcpyAsync(dy, yi.data(), yi.size() * sizeof(int4), hipMemcpyHostToDevice, st);
    hipMemcpyAsync(dx, xi.data(), xi.size() * sizeof(int4), hipMemcpyHostToDevice, st);
    const size_t sm = (size_t)(2 * T + 1) * maxK * sizeof(R2);
    // plane passes (as the engine uses them) when the planes fit
    const bool plane = fft_plane_enabled() && plane_lds_bytes<R2>(ny, nz) <= kMaxBlockLds &&
                       xpass_lds_bytes<R2>(nx, nzh) <= kMaxBlockLds;
    std::vector<int4> pi, bi;
    int4 *dp = nullptr, *db = nullptr;
    const size_t pl = plane_lds_bytes<R2>(ny, nz), xl = xpass_lds_bytes<R2>(nx, nzh);
    if (plane) {
        for (int x = 0; x < nx; ++x) pi.push_back(make_int4(0, x, 0, 0));
        for (int y = 0; y < ny; ++y) bi.push_back(make_int4(0, y, 0, nzh));
        hipMalloc(&dp, pi.size() * sizeof(int4));
        hipMalloc(&db, bi.size() * sizeof(int4));
        hipMemcpyAsync(dp, pi.data(), pi.size() * sizeof(int4), hipMemcpyHostToDevice, st);
        hipMemcpyAsync(db, bi.data(), bi.size() * sizeof(int4), hipMemcpyHostToDevice, st);
        hipFuncSetAttribute((const void*)fft_plane_fwd<P>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl);
        hipFuncSetAttribute((const void*)fft_plane_inv<P>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl);
        hipFuncSetAttribute((const void*)fft_x_block<P, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)xl);
        hipFuncSetAttribute((const void*)fft_x_block<P, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)xl);
    }
    if (forward) {
        std::vector<typename P::gacc> h(G);
        for (size_t i = 0; i < G; ++i) {
            h[i] = (long long)llrint(in[i] * P::grid_scale);
        }
        hipMemcpyAsync(rg, h.data(), G * sizeof(typename P::gacc), hipMemcpyHostToDevice, st);
        if (plane) {
            fft_plane_fwd<P><<<(unsigned)pi.size(), kPlaneThreads, pl, st>>>(dd, dp, rg, cg, dtw);
            fft_x_block<P, 0><<<(unsigned)bi.size(), kXPassThreads, xl, st>>>(dd, db, cg, nullptr, dtw, nullptr);
        } else {
            fft_z_fwd<P><<<(unsigned)zi.size(), 256, sm, st>>>(dd, dz, rg, cg, dtw);
            fft_y<P><<<(unsigned)yi.size(), 256, sm, st>>>(dd, dy, cg, dtw, T, -1);
            fft_x_plain<P><<<(unsigned)xi.size(), 256, sm, st>>>(dd, dx, cg, dtw, T, -1);
        }
        std::vector<R2> hc(Gh);
        hipMemcpyAsync(hc.data(), cg, Gh * sizeof(R2), hipMemcpyDeviceToHost, st);
        hipStreamSynchronize(st);
        for (size_t i = 0; i < Gh; ++i) { out[2 * i] = hc[i].x; out[2 * i + 1] = hc[i].y; }
    } else {
        std::vector<R2> hc(Gh);
        for (size_t i = 0; i < Gh; ++i) { hc[i].x = (real)in[2 * i]; hc[i].y = (real)in[2 * i + 1]; }
        hipMemcpyAsync(cg, hc.data(), Gh * sizeof(R2), hipMemcpyHostToDevice, st);
        if (plane) {
            fft_x_block<P, 1><<<(unsigned)bi.size(), kXPassThreads, xl, st>>>(dd, db, cg, nullptr, dtw, nullptr);
            fft_plane_inv<P><<<(unsigned)pi.size(), kPlaneThreads, pl, st>>>(dd, dp, cg, ph, dtw);
        } else {
            fft_x_plain<P><<<(unsigned)xi.size(), 256, sm, st>>>(dd, dx, cg, dtw, T, +1);
            fft_y<P><<<(unsigned)yi.size(), 256, sm, st>>>(dd, dy, cg, dtw, T, +1);
            fft_z_inv<P><<<(unsigned)zi.size(), 256, sm, st>>>(dd, dz, cg, ph, dtw);
        }
        std::vector<real> hr(G);
        hipMemcpyAsync(hr.data(), ph, G * sizeof(real), hipMemcpyDeviceToHost, st);
        hipStreamSynchronize(st);
        for (size_t i = 0; i < G; ++i) out[i] = hr[i];
    }
    hipFree(dd); hipFree(dtw); hipFree(dz); hipFree(dy); hipFree(dx); hipFree(rg); hipFree(cg); hipFree(ph);
    if (dp) hipFree(dp);
    if (db) hipFree(db);
}

template void fft3d_roundtrip<PrecFP64>(int, int, int, const double*, double*, bool, hipStream_t);
template void fft3d_roundtrip<PrecMixed>(int, int, int, const double*, double*, bool, hipStream_t);

// ------------------------------------------------------ no-allocation FFT
//
// The forward transform of fft3d_roundtrip with every device buffer carved
// out of one file-scope array, `scratch`, so that a call allocates nothing.
// Each GPU holds its own copy of the array from the library's load until the
// process ends.  Layout (256-byte aligned regions):
//   [SysDev][twiddles, 3 x kNoAllocMaxK][work items, kNoAllocItems int4]
//   [real grid, 8 B per point][half spectrum, sizeof(real2) per point]
// The input is copied straight into the real-grid region (the mixed path
// then converts it to fixed point in place, element by element), and the
// FP64 spectrum is copied straight out of its region; the mixed spectrum
// (float pairs) is copied into the caller's output and widened there in
// place, from the last element backwards.  The host side uses the device
// address returned by hipGetSymbolAddress, never the array's name.
constexpr int kNoAllocMaxK = 256;
constexpr long long kNoAllocMaxPoints = 1LL << 20;
constexpr int kNoAllocItems = 32768;
constexpr size_t kNoAllocBytes = 20u << 20;
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

std::mutex g_noalloc_mu[kNoAllocMaxDevices];     // one call at a time per device
void* g_noalloc_ptr[kNoAllocMaxDevices] = {};    // device address of each device's array

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Work items of the forward passes, in fft3d_roundtrip's order.  Plane passes:
// x planes, then y rows of the x pass over every kz.  Line passes: groups of
// T z lines, then (x, kz block) and (y, kz block) items.
__global__ void noalloc_items_kernel(int4* it, int nx, int ny, int nzh, int T, int plane, int n0, int n1, int n2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n0 + n1 + n2) return;
    if (plane) {
        if (i < n0) it[i] = make_int4(0, i, 0, 0);
        else it[i] = make_int4(0, i - n0, 0, nzh);
        return;
    }
    const int nkb = (nzh + T - 1) / T;
    if (i < n0) {
        const int l = i * T;
        it[i] = make_int4(0, l, min(T, nx * ny - l), 0);
    } else if (i < n0 + n1) {
        const int j = i - n0;
        it[i] = make_int4(0, j / nkb, (j % nkb) * T, 0);
    } else {
        const int j = i - n0 - n1;
        it[i] = make_int4(0, j / nkb, (j % nkb) * T, 0);
    }
}

// In place: the doubles of the input become 2^32 fixed point (as the host
// conversion of fft3d_roundtrip: round to nearest, ties to even).
__global__ void noalloc_to_fixed_kernel(void* g, long long n, double scale) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = reinterpret_cast<const double*>(g)[i];
    reinterpret_cast<long long*>(g)[i] = __double2ll_rn(v * scale);
}

}  // namespace

template <class P>
nx_status fft3d_r2c_noalloc(int device, int nx, int ny, int nz, const double* in, double* out, std::string& err) {
    using R2 = typename P::real2;
    using real = typename P::real;
    if (device < 0 || device >= kNoAllocMaxDevices) { err = "device index out of range"; return NX_ERR_ARG; }
    const int nzh = nz / 2 + 1;
    const long long G = (long long)nx * ny * nz, Gh = (long long)nx * ny * nzh;
    if (nx > kNoAllocMaxK || ny > kNoAllocMaxK || nz > kNoAllocMaxK || G > kNoAllocMaxPoints) {
        err = "grid " + std::to_string(nx) + "x" + std::to_string(ny) + "x" + std::to_string(nz) +
              " exceeds the no-allocation maximum (each dimension <= " + std::to_string(kNoAllocMaxK) + ", at most " +
              std::to_string(kNoAllocMaxPoints) + " points)";
        return NX_ERR_CAPACITY;
    }
    const int maxK = std::max(nx, std::max(ny, nz));
    const int T = std::max(1, std::min(8, (int)((61440 / (int)sizeof(R2) - maxK) / (2 * maxK))));
    const bool plane = fft_plane_enabled() && plane_lds_bytes<R2>(ny, nz) <= kMaxBlockLds &&
                       xpass_lds_bytes<R2>(nx, nzh) <= kMaxBlockLds;
    const int nkb = (nzh + T - 1) / T;
    const int n0 = plane ? nx : (nx * ny + T - 1) / T;
    const int n1 = plane ? ny : nx * nkb;
    const int n2 = plane ? 0 : ny * nkb;
    const size_t o_tw = align256(sizeof(SysDev));
    const size_t o_it = o_tw + align256(3 * kNoAllocMaxK * sizeof(R2));
    const size_t o_rg = o_it + align256((size_t)kNoAllocItems * sizeof(int4));
    const size_t o_cg = o_rg + align256((size_t)G * 8);
    const size_t need = o_cg + (size_t)Gh * sizeof(R2);
    if (n0 + n1 + n2 > kNoAllocItems || need > kNoAllocBytes) {
        err = "grid " + std::to_string(nx) + "x" + std::to_string(ny) + "x" + std::to_string(nz) + " needs " +
              std::to_string(need) + " bytes of scratch; the no-allocation array holds " + std::to_string(kNoAllocBytes);
        return NX_ERR_CAPACITY;
    }
    std::lock_guard<std::mutex> lock(g_noalloc_mu[device]);
    auto fail = [&](hipError_t e, const char* what) {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return NX_ERR_HIP;
    };
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return fail(e, "hipSetDevice");
    if (!g_noalloc_ptr[device]) {
        // once per device, with that device current: the address of its copy
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail(e, "hipGetSymbolAddress(scratch)");
        g_noalloc_ptr[device] = p;
    }
    unsigned char* base = static_cast<unsigned char*>(g_noalloc_ptr[device]);
    SysDev* dd = reinterpret_cast<SysDev*>(base);
    R2* dtw = reinterpret_cast<R2*>(base + o_tw);
    int4* items = reinterpret_cast<int4*>(base + o_it);
    typename P::gacc* rg = reinterpret_cast<typename P::gacc*>(base + o_rg);
    R2* cg = reinterpret_cast<R2*>(base + o_cg);
    const hipStream_t st = nullptr;
    // the per-call tables, from host values on the stack
    SysDev d{};
    d.grid[0] = nx; d.grid[1] = ny; d.grid[2] = nz;
    d.tw_off[0] = 0; d.tw_off[1] = nx; d.tw_off[2] = nx + ny;
    R2 tw[3 * kNoAllocMaxK];
    const int K[3] = {nx, ny, nz};
    for (int a = 0, o = 0; a < 3; o += K[a], ++a)
        for (int t = 0; t < K[a]; ++t) {
            const double ang = -2.0 * kPi * (double)t / (double)K[a];
            tw[o + t].x = (real)cos(ang);
            tw[o + t].y = (real)sin(ang);
        }
    if ((e = hipMemcpyAsync(dd, &d, sizeof(SysDev), hipMemcpyHostToDevice, st)) != hipSuccess) return fail(e, "copy");
    if ((e = hipMemcpyAsync(dtw, tw, (size_t)(nx + ny + nz) * sizeof(R2), hipMemcpyHostToDevice, st)) != hipSuccess)
        return fail(e, "copy");
    noalloc_items_kernel<<<(n0 + n1 + n2 + 255) / 256, 256, 0, st>>>(items, nx, ny, nzh, T, plane ? 1 : 0, n0, n1, n2);
    if ((e = hipMemcpyAsync(rg, in, (size_t)G * sizeof(double), hipMemcpyHostToDevice, st)) != hipSuccess)
        return fail(e, "copy");
    noalloc_to_fixed_kernel<<<(unsigned)((G + 255) / 256), 256, 0, st>>>(rg, G, P::grid_scale);
    if (plane) {
        const size_t pl = plane_lds_bytes<R2>(ny, nz), xl = xpass_lds_bytes<R2>(nx, nzh);
        if ((e = hipFuncSetAttribute((const void*)fft_plane_fwd<P>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)pl)) != hipSuccess ||
            (e = hipFuncSetAttribute((const void*)fft_x_block<P, 0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)xl)) != hipSuccess)
            return fail(e, "hipFuncSetAttribute");
        fft_plane_fwd<P><<<(unsigned)n0, kPlaneThreads, pl, st>>>(dd, items, rg, cg, dtw);
        fft_x_block<P, 0><<<(unsigned)n1, kXPassThreads, xl, st>>>(dd, items + n0, cg, nullptr, dtw, nullptr);
    } else {
        const size_t sm = (size_t)(2 * T + 1) * maxK * sizeof(R2);
        fft_z_fwd<P><<<(unsigned)n0, 256, sm, st>>>(dd, items, rg, cg, dtw);
        fft_y<P><<<(unsigned)n1, 256, sm, st>>>(dd, items + n0, cg, dtw, T, -1);
        fft_x_plain<P><<<(unsigned)n2, 256, sm, st>>>(dd, items + n0 + n1, cg, dtw, T, -1);
    }
    if ((e = hipGetLastError()) != hipSuccess) return fail(e, "kernel launch");
    if ((e = hipMemcpyAsync(out, cg, (size_t)Gh * sizeof(R2), hipMemcpyDeviceToHost, st)) != hipSuccess)
        return fail(e, "copy");
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail(e, "hipStreamSynchronize");
    if constexpr (P::mixed) {
        // float pairs at the front of `out` widened to doubles in place, last first:
        // element i is read from bytes [8i, 8i + 8) before bytes [16i, 16i + 16) are
        // written, and no element below i is read from at or above byte 8i
        unsigned char* ob = reinterpret_cast<unsigned char*>(out);
        for (long long i = Gh - 1; i >= 0; --i) {
            float v[2];
            std::memcpy(v, ob + 8 * i, sizeof(v));
            const double w[2] = {(double)v[0], (double)v[1]};
            std::memcpy(ob + 16 * i, w, sizeof(w));
        }
    }
    return NX_OK;
}

template nx_status fft3d_r2c_noalloc<PrecFP64>(int, int, int, int, const double*, double*, std::string&);
template nx_status fft3d_r2c_noalloc<PrecMixed>(int, int, int, int, const double*, double*, std::string&);

}  // namespace nx

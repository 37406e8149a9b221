// bonded.hip -- harmonic bond, harmonic angle and periodic torsion kernels.
//
// One thread per term; forces are added to the atom-order accumulators.
// Formulas and operation order follow the reference:
//   bonds    core/baseline/src/Forces/HarmonicBondForce.cpp:20-58
//   angles   core/baseline/src/Forces/HarmonicAngleForce.cpp:19-63
//   torsions core/baseline/src/Forces/PeriodicTorsionForce.cpp:41-149
// Bond vectors use the reference's single-correction minimum image on the
// FP64 positions; the mixed path then evaluates the term in FP32.
#include "nx_device.h"
#include "nx_engine.h"

namespace nx {

template <typename real>
struct V3 { real x, y, z; };

template <typename real>
__device__ __forceinline__ V3<real> cross(V3<real> a, V3<real> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename real>
__device__ __forceinline__ real dot(V3<real> a, V3<real> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

template <typename real>
__device__ __forceinline__ V3<real> mi(const SysDev& d, const double* X, const double* Y, const double* Z, int a, int b) {
    // a - b with one +-L correction (PeriodicBoundaryCondition.cpp:93-104)
    return {(real)min_image1(X[a] - X[b], d.L[0]), (real)min_image1(Y[a] - Y[b], d.L[1]),
            (real)min_image1(Z[a] - Z[b], d.L[2])};
}

template <class P>
__device__ __forceinline__ void add3(typename P::facc* FX, typename P::facc* FY, typename P::facc* FZ, int a,
                                     V3<typename P::real> f) {
    facc_add(&FX[a], f.x);
    facc_add(&FY[a], f.y);
    facc_add(&FZ[a], f.z);
}

template <class P, bool ENERGY>
__global__ void bond_kernel(const SysDev* __restrict__ sd, const int* __restrict__ atom_sys,
                            const int2* __restrict__ idx, const double2* __restrict__ par, int n,
                            const double* __restrict__ X, const double* __restrict__ Y, const double* __restrict__ Z,
                            typename P::facc* FX, typename P::facc* FY, typename P::facc* FZ, long long* energy) {
    using real = typename P::real;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = t < n;
    int s = 0;
    double eterm = 0.0;
    if (valid) {
    const int2 ij = idx[t];
    s = atom_sys[ij.x];
    const SysDev& d = sd[s];
    V3<real> dl = mi<real>(d, X, Y, Z, ij.y, ij.x);   // pos2 - pos1
    const double2 pk = par[t];
    const real r = sqrt(dot(dl, dl));
    const real dI = r - (real)pk.x;
    const real dEdR = (real)pk.y * dI;
    V3<real> f1 = {0, 0, 0};
    if (r > 0) {
        const real sc = dEdR / r;
        f1 = {dl.x * sc, dl.y * sc, dl.z * sc};
    }
    add3<P>(FX, FY, FZ, ij.x, f1);
    add3<P>(FX, FY, FZ, ij.y, V3<real>{-f1.x, -f1.y, -f1.z});
    eterm = 0.5 * pk.y * (double)dI * (double)dI;
    }
    if constexpr (ENERGY) wave_energy_add(energy, s, E_BOND, eterm, valid);
}

template <class P, bool ENERGY>
__global__ void angle_kernel(const SysDev* __restrict__ sd, const int* __restrict__ atom_sys,
                             const int4* __restrict__ idx, const double2* __restrict__ par, int n,
                             const double* __restrict__ X, const double* __restrict__ Y, const double* __restrict__ Z,
                             typename P::facc* FX, typename P::facc* FY, typename P::facc* FZ, long long* energy) {
    using real = typename P::real;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = t < n;
    int s = 0;
    double eterm = 0.0;
    if (valid) {
    const int4 ijk = idx[t];
    s = atom_sys[ijk.x];
    const SysDev& d = sd[s];
    const V3<real> v0 = mi<real>(d, X, Y, Z, ijk.y, ijk.x);   // p2 - p1
    const V3<real> v1 = mi<real>(d, X, Y, Z, ijk.y, ijk.z);   // p2 - p3
    const double2 pk = par[t];
    const real r21 = sqrt(dot(v0, v0)), r23 = sqrt(dot(v1, v1));
    real c = dot(v0, v1) / (r21 * r23);
    c = c < (real)-1 ? (real)-1 : (c > (real)1 ? (real)1 : c);
    const real theta = acos(c);
    const real dI = theta - (real)pk.x;
    const real dE = (real)pk.y * dI;
    const V3<real> cp = cross(v0, v1);
    real rp = sqrt(dot(cp, cp));
    rp = rp < (real)1.0e-6 ? (real)1.0e-6 : rp;
    const real q21 = dot(v0, v0), q23 = dot(v1, v1);
    V3<real> f1 = cross(v0, cp), f3 = cross(cp, v1);
    const real s1 = dE / (q21 * rp), s3 = dE / (q23 * rp);
    f1 = {f1.x * s1, f1.y * s1, f1.z * s1};
    f3 = {f3.x * s3, f3.y * s3, f3.z * s3};
    const V3<real> f2 = {-f1.x - f3.x, -f1.y - f3.y, -f1.z - f3.z};
    add3<P>(FX, FY, FZ, ijk.x, f1);
    add3<P>(FX, FY, FZ, ijk.y, f2);
    add3<P>(FX, FY, FZ, ijk.z, f3);
    eterm = 0.5 * pk.y * (double)dI * (double)dI;
    }
    if constexpr (ENERGY) wave_energy_add(energy, s, E_ANGLE, eterm, valid);
}

template <typename real>
__device__ __forceinline__ V3<real> unit(V3<real> v) {
    const real len = sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    if (len > 0) return {v.x / len, v.y / len, v.z / len};
    return {0, 0, 0};
}

template <class P, bool ENERGY>
__global__ void torsion_kernel(const SysDev* __restrict__ sd, const int* __restrict__ atom_sys,
                               const int4* __restrict__ idx, const double4* __restrict__ par, int n,
                               const double* __restrict__ X, const double* __restrict__ Y,
                               const double* __restrict__ Z, typename P::facc* FX, typename P::facc* FY,
                               typename P::facc* FZ, long long* energy) {
    using real = typename P::real;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = t < n;
    int s = 0;
    double eterm = 0.0;
    if (valid) {
    const int4 q = idx[t];
    s = atom_sys[q.x];
    const SysDev& d = sd[s];
    const V3<real> v0 = mi<real>(d, X, Y, Z, q.x, q.y);   // r1 - r2
    const V3<real> v1 = mi<real>(d, X, Y, Z, q.z, q.y);   // r3 - r2
    const V3<real> v2 = mi<real>(d, X, Y, Z, q.z, q.w);   // r3 - r4
    const double4 pk = par[t];
    const V3<real> cp0 = cross(v0, v1), cp1 = cross(v1, v2);
    const real cosang = dot(unit(cp0), unit(cp1));
    real ang;
    const real thr = (real)(double)0.99f;   // the reference compares against the float literal 0.99f
    if (cosang > thr || cosang < -thr) {
        const V3<real> cr = cross(cp0, cp1);
        const real scale = dot(cp0, cp0) * dot(cp1, cp1);
        ang = asin(sqrt(dot(cr, cr) / scale));
        if (cosang < 0) ang = (real)kPi - ang;
    } else {
        ang = acos(cosang);
    }
    ang = (dot(v0, cp1) >= 0 ? ang : -ang);
    const real per = (real)pk.z;
    const real delta = per * ang - (real)pk.y;
    const real dEdPhi = -(real)pk.x * per * sin(delta);
    const real nc1 = dot(cp0, cp0), nsq = dot(v1, v1), nbc = sqrt(nsq), nc2 = dot(cp1, cp1);
    const real dp = (real)1.0 / nsq;
    const real fa1 = (-dEdPhi * nbc) / nc1;
    const real fa2 = dot(v0, v1) * dp;
    const real fa3 = dot(v2, v1) * dp;
    const real fa4 = (dEdPhi * nbc) / nc2;
    const V3<real> F1 = {cp0.x * fa1, cp0.y * fa1, cp0.z * fa1};
    const V3<real> F4 = {cp1.x * fa4, cp1.y * fa4, cp1.z * fa4};
    const V3<real> sv = {F1.x * fa2 - F4.x * fa3, F1.y * fa2 - F4.y * fa3, F1.z * fa2 - F4.z * fa3};
    add3<P>(FX, FY, FZ, q.x, F1);
    add3<P>(FX, FY, FZ, q.y, V3<real>{sv.x - F1.x, sv.y - F1.y, sv.z - F1.z});
    add3<P>(FX, FY, FZ, q.z, V3<real>{-sv.x - F4.x, -sv.y - F4.y, -sv.z - F4.z});
    add3<P>(FX, FY, FZ, q.w, F4);
    eterm = pk.x * (1.0 + cos((double)delta));
    }
    if constexpr (ENERGY) wave_energy_add(energy, s, E_TORSION, eterm, valid);
}

template <class P>
void launch_bonded(Batch<P>& B, int mask, bool energy) {
    const int T = 256;
    auto nb = [](int n) { return (unsigned)((n + 255) / 256); };
    if ((mask & 1) && B.n_bonds) {
        if (energy) bond_kernel<P, true><<<nb(B.n_bonds), T, 0, B.stream>>>(B.d_sys, B.atom_sys, B.bond_idx, B.bond_par, B.n_bonds, B.X, B.Y, B.Z, B.FAx, B.FAy, B.FAz, B.energy);
        else bond_kernel<P, false><<<nb(B.n_bonds), T, 0, B.stream>>>(B.d_sys, B.atom_sys, B.bond_idx, B.bond_par, B.n_bonds, B.X, B.Y, B.Z, B.FAx, B.FAy, B.FAz, B.energy);
    }
    if ((mask & 2) && B.n_angles) {
        if (energy) angle_kernel<P, true><<<nb(B.n_angles), T, 0, B.stream>>>(B.d_sys, B.atom_sys, B.angle_idx, B.angle_par, B.n_angles, B.X, B.Y, B.Z, B.FAx, B.FAy, B.FAz, B.energy);
        else angle_kernel<P, false><<<nb(B.n_angles), T, 0, B.stream>>>(B.d_sys, B.atom_sys, B.angle_idx, B.angle_par, B.n_angles, B.X, B.Y, B.Z, B.FAx, B.FAy, B.FAz, B.energy);
    }
    if ((mask & 4) && B.n_tors) {
        if (energy) torsion_kernel<P, true><<<nb(B.n_tors), T, 0, B.stream>>>(B.d_sys, B.atom_sys, B.tors_idx, B.tors_par, B.n_tors, B.X, B.Y, B.Z, B.FAx, B.FAy, B.FAz, B.energy);
        else torsion_kernel<P, false><<<nb(B.n_tors), T, 0, B.stream>>>(B.d_sys, B.atom_sys, B.tors_idx, B.tors_par, B.n_tors, B.X, B.Y, B.Z, B.FAx, B.FAy, B.FAz, B.energy);
    }
}

template void launch_bonded<PrecFP64>(Batch<PrecFP64>&, int, bool);
template void launch_bonded<PrecMixed>(Batch<PrecMixed>&, int, bool);

}  // namespace nx

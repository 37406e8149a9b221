// bonded.hip -- harmonic bond, harmonic angle and periodic torsion terms, and
// the per-atom force gather.
//
// bonded_kernel: one thread per term (bonds, then angles, then torsions, in
// one launch).  Each term writes the force on each of its atoms into its own
// slots of the per-term force buffer tF (plain stores).
// atom_force_kernel: one thread per atom sums the term forces acting on it
// (per-atom CSR of slots, fixed order) and its exclusion corrections, and
// stores the atom-order force.  No force is added atomically.
//
// Formulas and operation order follow the reference:
//   bonds      core/baseline/src/Forces/HarmonicBondForce.cpp:20-58
//   angles     core/baseline/src/Forces/HarmonicAngleForce.cpp:19-63
//   torsions   core/baseline/src/Forces/PeriodicTorsionForce.cpp:41-149
//   exclusions core/baseline/src/Forces/NonbondedForce.cpp:590-630
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

template <typename real>
__device__ __forceinline__ void put(real* tF, int n, int slot, V3<real> f) {
    tF[slot] = f.x;
    tF[n + slot] = f.y;
    tF[2 * n + slot] = f.z;
}

template <typename real>
__device__ __forceinline__ V3<real> unit(V3<real> v) {
    const real len = sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    if (len > 0) return {v.x / len, v.y / len, v.z / len};
    return {0, 0, 0};
}

// Harmonic bond: forces on p1, p2 and the energy.
template <typename real>
__device__ __forceinline__ double bond_term(const SysDev& d, const double* X, const double* Y, const double* Z,
                                            int2 ij, double2 pk, V3<real>& f1, V3<real>& f2) {
    V3<real> dl = mi<real>(d, X, Y, Z, ij.y, ij.x);   // pos2 - pos1
    const real r = sqrt(dot(dl, dl));
    const real dI = r - (real)pk.x;
    const real dEdR = (real)pk.y * dI;
    f1 = {0, 0, 0};
    if (r > 0) {
        const real sc = dEdR / r;
        f1 = {dl.x * sc, dl.y * sc, dl.z * sc};
    }
    f2 = {-f1.x, -f1.y, -f1.z};
    return 0.5 * pk.y * (double)dI * (double)dI;
}

// Harmonic angle: forces on p1, p2, p3 and the energy.
template <typename real>
__device__ __forceinline__ double angle_term(const SysDev& d, const double* X, const double* Y, const double* Z,
                                             int4 ijk, double2 pk, V3<real>& f1, V3<real>& f2, V3<real>& f3) {
    const V3<real> v0 = mi<real>(d, X, Y, Z, ijk.y, ijk.x);   // p2 - p1
    const V3<real> v1 = mi<real>(d, X, Y, Z, ijk.y, ijk.z);   // p2 - p3
    const V3<real> cp = cross(v0, v1);
    real dI;
    if constexpr (sizeof(real) == 8) {
        // the reference's arithmetic: acos of the clamped cosine
        const real r21 = sqrt(dot(v0, v0)), r23 = sqrt(dot(v1, v1));
        real c = dot(v0, v1) / (r21 * r23);
        c = c < (real)-1 ? (real)-1 : (c > (real)1 ? (real)1 : c);
        dI = acos(c) - (real)pk.x;
    } else {
        // FP32: the angle from atan2(|v0 x v1|, v0 . v1), accurate at every angle
        // (acos of a cosine near +-1 loses half the digits); an obtuse angle is
        // measured from pi, so that theta - theta0 near a straight angle is the
        // difference of two small, accurate numbers
        const real sn = sqrt(dot(cp, cp)), cs = dot(v0, v1);
        if (cs < 0) dI = (real)(kPi - pk.x) - atan2(sn, -cs);
        else dI = atan2(sn, cs) - (real)pk.x;
    }
    const real dE = (real)pk.y * dI;
    real rp = sqrt(dot(cp, cp));
    rp = rp < (real)1.0e-6 ? (real)1.0e-6 : rp;
    const real q21 = dot(v0, v0), q23 = dot(v1, v1);
    f1 = cross(v0, cp);
    f3 = cross(cp, v1);
    const real s1 = dE / (q21 * rp), s3 = dE / (q23 * rp);
    f1 = {f1.x * s1, f1.y * s1, f1.z * s1};
    f3 = {f3.x * s3, f3.y * s3, f3.z * s3};
    f2 = {-f1.x - f3.x, -f1.y - f3.y, -f1.z - f3.z};
    return 0.5 * pk.y * (double)dI * (double)dI;
}

// Periodic torsion: forces on p1..p4 and the energy.
template <typename real>
__device__ __forceinline__ double torsion_term(const SysDev& d, const double* X, const double* Y, const double* Z,
                                               int4 q, double4 pk, V3<real>& F1, V3<real>& F2, V3<real>& F3,
                                               V3<real>& F4) {
    const V3<real> v0 = mi<real>(d, X, Y, Z, q.x, q.y);   // r1 - r2
    const V3<real> v1 = mi<real>(d, X, Y, Z, q.z, q.y);   // r3 - r2
    const V3<real> v2 = mi<real>(d, X, Y, Z, q.z, q.w);   // r3 - r4
    const V3<real> cp0 = cross(v0, v1), cp1 = cross(v1, v2);
    real ang;
    if constexpr (sizeof(real) == 8) {
        // the reference's arithmetic: acos of the cosine, asin where |cos| > 0.99
        const real cosang = dot(unit(cp0), unit(cp1));
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
    } else {
        // FP32: phi = atan2(|v1| v0 . cp1, cp0 . cp1), accurate at every angle.  The
        // same angle and sign: cp0 x cp1 = v1 (v0 . cp1), so |v1| |v0 . cp1| is
        // |cp0 x cp1|, and the sign is the reference's sign of v0 . cp1
        ang = atan2(sqrt(dot(v1, v1)) * dot(v0, cp1), dot(cp0, cp1));
    }
    const real per = (real)pk.z;
    const real delta = per * ang - (real)pk.y;
    const real dEdPhi = -(real)pk.x * per * sin(delta);
    const real nc1 = dot(cp0, cp0), nsq = dot(v1, v1), nbc = sqrt(nsq), nc2 = dot(cp1, cp1);
    const real dp = (real)1.0 / nsq;
    const real fa1 = (-dEdPhi * nbc) / nc1;
    const real fa2 = dot(v0, v1) * dp;
    const real fa3 = dot(v2, v1) * dp;
    const real fa4 = (dEdPhi * nbc) / nc2;
    F1 = {cp0.x * fa1, cp0.y * fa1, cp0.z * fa1};
    F4 = {cp1.x * fa4, cp1.y * fa4, cp1.z * fa4};
    const V3<real> sv = {F1.x * fa2 - F4.x * fa3, F1.y * fa2 - F4.y * fa3, F1.z * fa2 - F4.z * fa3};
    F2 = {sv.x - F1.x, sv.y - F1.y, sv.z - F1.z};
    F3 = {-sv.x - F4.x, -sv.y - F4.y, -sv.z - F4.z};
    return pk.x * (1.0 + cos((double)delta));
}

// Term t of [0, nb + na + nt): bonds, then angles, then torsions.  Terms of a
// force that is not enabled store zero forces.
template <class P, bool ENERGY>
__global__ void bonded_kernel(const SysDev* __restrict__ sd, const int* __restrict__ atom_sys,
                              const int2* __restrict__ bidx, const double2* __restrict__ bpar, int nb,
                              const int4* __restrict__ aidx, const double2* __restrict__ apar, int na,
                              const int4* __restrict__ tidx, const double4* __restrict__ tpar, int nt,
                              const double* __restrict__ X, const double* __restrict__ Y,
                              const double* __restrict__ Z, typename P::real* tF, int n_tslots, int mask,
                              long long* energy) {
    side_priority();
    using real = typename P::real;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int s = 0, kind = -1;
    double eterm = 0.0;
    const V3<real> zero = {0, 0, 0};
    if (t < nb) {
        kind = 0;
        const int2 ij = bidx[t];
        s = atom_sys[ij.x];
        V3<real> f1 = zero, f2 = zero;
        if (mask & 1) eterm = bond_term<real>(sd[s], X, Y, Z, ij, bpar[t], f1, f2);
        put(tF, n_tslots, 2 * t, f1);
        put(tF, n_tslots, 2 * t + 1, f2);
    } else if (t < nb + na) {
        kind = 1;
        const int u = t - nb;
        const int4 ijk = aidx[u];
        s = atom_sys[ijk.x];
        V3<real> f1 = zero, f2 = zero, f3 = zero;
        if (mask & 2) eterm = angle_term<real>(sd[s], X, Y, Z, ijk, apar[u], f1, f2, f3);
        const int b = 2 * nb + 3 * u;
        put(tF, n_tslots, b, f1);
        put(tF, n_tslots, b + 1, f2);
        put(tF, n_tslots, b + 2, f3);
    } else if (t < nb + na + nt) {
        kind = 2;
        const int u = t - nb - na;
        const int4 q = tidx[u];
        s = atom_sys[q.x];
        V3<real> f1 = zero, f2 = zero, f3 = zero, f4 = zero;
        if (mask & 4) eterm = torsion_term<real>(sd[s], X, Y, Z, q, tpar[u], f1, f2, f3, f4);
        const int b = 2 * nb + 3 * na + 4 * u;
        put(tF, n_tslots, b, f1);
        put(tF, n_tslots, b + 1, f2);
        put(tF, n_tslots, b + 2, f3);
        put(tF, n_tslots, b + 3, f4);
    }
    if constexpr (ENERGY) {
        wave_energy_add(energy, s, E_BOND, eterm, kind == 0);
        wave_energy_add(energy, s, E_ANGLE, eterm, kind == 1);
        wave_energy_add(energy, s, E_TORSION, eterm, kind == 2);
    }
}

// Per atom: the sum of its term forces (slot order of the CSR, fixed) and, with
// the nonbonded force on, its exclusion corrections, computed from each side
// of the pair over the symmetric exclusion lists (the energy of a pair is
// counted from its lower-index atom).  Stores the atom-order force.
template <class P, bool ENERGY>
__global__ void atom_force_kernel(const SysDev* __restrict__ sd, const int* __restrict__ atom_sys,
                                  const int* __restrict__ at_off, const int* __restrict__ at_idx,
                                  const typename P::real* __restrict__ tF, int n_tslots,
                                  const int* __restrict__ excl_off, const int* __restrict__ excl_idx,
                                  const double* __restrict__ X, const double* __restrict__ Y,
                                  const double* __restrict__ Z, const double* __restrict__ Q, typename P::facc* FX,
                                  typename P::facc* FY, typename P::facc* FZ, long long* energy, int n_atoms,
                                  int mask) {
    side_priority();
    using real = typename P::real;
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    double eacc = 0.0;
    int s = 0;
    const bool valid = a < n_atoms;
    if (valid) {
        s = atom_sys[a];
        real fx = 0, fy = 0, fz = 0;
        if (mask & 7) {
            const int b1 = at_off[a + 1];
            for (int b = at_off[a]; b < b1; ++b) {
                const int k = at_idx[b];
                fx += tF[k];
                fy += tF[n_tslots + k];
                fz += tF[2 * n_tslots + k];
            }
        }
        if (mask & NX_FORCE_NONBONDED_BIT) {
            const SysDev& d = sd[s];
            const double xa = X[a], ya = Y[a], za = Z[a], qa = kOne4PiEps0 * Q[a];
            const int e1 = excl_off[a + 1];
            for (int e = excl_off[a]; e < e1; ++e) {
                const int j = excl_idx[e];
                const double dxd = min_image1(xa - X[j], d.L[0]);
                const double dyd = min_image1(ya - Y[j], d.L[1]);
                const double dzd = min_image1(za - Z[j], d.L[2]);
                const double qq = qa * Q[j];
                double ej;
                if constexpr (P::mixed) {
                    const float dx = (float)dxd, dy = (float)dyd, dz = (float)dzd;
                    const float r = sqrtf(dx * dx + dy * dy + dz * dz);
                    const float ir = 1.0f / r;
                    const float ar = (float)d.alpha * r;
                    const float ef = erff(ar);
                    const float qf = (float)qq;
                    if (ef > 1e-6f) {
                        const float dEdR =
                            qf * ir * ir * ir * (ef - 2.0f * ar * __expf(-ar * ar) * (float)(1.0 / kSqrtPi));
                        fx -= dx * dEdR; fy -= dy * dEdR; fz -= dz * dEdR;
                        ej = (double)(qf * ir * ef);
                    } else {
                        ej = d.alpha * kTwoOverSqrtPi * qq;
                    }
                } else {
                    const double r = sqrt(dxd * dxd + dyd * dyd + dzd * dzd);
                    const double ir = 1.0 / r;
                    const double ar = d.alpha * r;
                    const double ef = erf(ar);
                    if (ef > 1e-6) {
                        const double dEdR = qq * ir * ir * ir * (ef - 2.0 * ar * exp(-ar * ar) / kSqrtPi);
                        fx -= dxd * dEdR; fy -= dyd * dEdR; fz -= dzd * dEdR;
                        ej = qq * ir * ef;
                    } else {
                        ej = d.alpha * kTwoOverSqrtPi * qq;
                    }
                }
                if (j > a) eacc += ej;
            }
        }
        facc_set(&FX[a], fx);
        facc_set(&FY[a], fy);
        facc_set(&FZ[a], fz);
    }
    if constexpr (ENERGY) wave_energy_add(energy, s, E_EXCL, eacc, valid && (mask & NX_FORCE_NONBONDED_BIT));
}

template <class P>
void launch_bonded(Batch<P>& B, int mask, bool energy) {
    const int n = B.n_bonds + B.n_angles + B.n_tors;
    if (n == 0 || !(mask & 7)) return;
    const int T = 256;
    const unsigned g = (unsigned)((n + T - 1) / T);
    if (energy)
        bonded_kernel<P, true><<<g, T, 0, B.stream>>>(B.d_sys, B.atom_sys, B.bond_idx, B.bond_par, B.n_bonds,
                                                      B.angle_idx, B.angle_par, B.n_angles, B.tors_idx, B.tors_par,
                                                      B.n_tors, B.X, B.Y, B.Z, B.tF, B.n_tslots, mask, B.energy);
    else
        bonded_kernel<P, false><<<g, T, 0, B.stream>>>(B.d_sys, B.atom_sys, B.bond_idx, B.bond_par, B.n_bonds,
                                                       B.angle_idx, B.angle_par, B.n_angles, B.tors_idx, B.tors_par,
                                                       B.n_tors, B.X, B.Y, B.Z, B.tF, B.n_tslots, mask, B.energy);
}

template <class P>
void launch_atom_forces(Batch<P>& B, int mask, bool energy) {
    const int T = 256;
    const unsigned g = (unsigned)((B.n_atoms + T - 1) / T);
    const int m = mask & ((B.n_tslots ? 7 : 0) | (B.n_excl_pairs ? NX_FORCE_NONBONDED_BIT : 0));
    if (energy)
        atom_force_kernel<P, true><<<g, T, 0, B.stream>>>(B.d_sys, B.atom_sys, B.at_off, B.at_idx, B.tF,
                                                          B.n_tslots, B.excl_off, B.excl_idx, B.X, B.Y, B.Z, B.q,
                                                          B.FAx, B.FAy, B.FAz, B.energy, B.n_atoms, m);
    else
        atom_force_kernel<P, false><<<g, T, 0, B.stream>>>(B.d_sys, B.atom_sys, B.at_off, B.at_idx, B.tF,
                                                           B.n_tslots, B.excl_off, B.excl_idx, B.X, B.Y, B.Z, B.q,
                                                           B.FAx, B.FAy, B.FAz, B.energy, B.n_atoms, m);
}

template void launch_bonded<PrecFP64>(Batch<PrecFP64>&, int, bool);
template void launch_bonded<PrecMixed>(Batch<PrecMixed>&, int, bool);
template void launch_atom_forces<PrecFP64>(Batch<PrecFP64>&, int, bool);
template void launch_atom_forces<PrecMixed>(Batch<PrecMixed>&, int, bool);

}  // namespace nx

// nonbonded.hip -- real-space LJ + erfc-Coulomb pair kernel, exclusion
// correction, and the explicit pair-set export.
//
// Physics follows NonbondedForce::DirectForcesAndEnergy and
// CalculateExclusionEnergyAndForces (core/baseline/src/Forces/
// NonbondedForce.cpp:530-585, :590-630) with the modified parameters of
// NonbondedParamsModifier (:504-513): sigma_ij = sig_i/2 + sig_j/2 and
// eps'_i = 2 sqrt(eps_i), so E_LJ = eps'_i eps'_j (s^6 - 1) s^6.
//
// Work decomposition: the list is cut into chunks of up to kChunkEntries
// entries of one super-cluster (up to 8 clusters of 8 atoms); the wavefronts
// of each XCD claim that XCD's chunks one at a time from a counter.  Lane
// l = 8*ia + ja holds i-atom ia of the current i-cluster and j-atom ja of the
// current j-cluster, so one (i-cluster, j-cluster) tile is 64 pairs in one
// VALU pass.  i-forces stay in registers over the chunk; j-forces are reduced
// across the 8 lanes sharing ja and added once per entry.  The FP64 path tests r^2 < rc^2 with the reference's
// minimum image and an uncontracted left-to-right dot product, so its pair
// set is bit-identical to the reference's.
#include <algorithm>

#include "nx_device.h"
#include "nx_engine.h"

namespace nx {

// Abramowitz & Stegun 7.1.26: erfc(x) ~ t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-x^2),
// t = 1 / (1 + p x), |error| <= 1.5e-7 for x >= 0.
constexpr double kAsP = 0.3275911;
__device__ __forceinline__ float erfc_as_t(float t, float expmx2) {
    const float a1 = 0.254829592f, a2 = -0.284496736f, a3 = 1.421413741f, a4 = -1.453152027f,
                a5 = 1.061405429f;
    return t * (a1 + t * (a2 + t * (a3 + t * (a4 + t * a5)))) * expmx2;
}

constexpr double kLog2E = 1.4426950408889634;   // log2(e)

// Per-system constants of the mixed pair function, formed once per chunk so
// that the per-pair work starts from r and r^2 directly.
struct MixedPair {
    float ap;       // p alpha: the A&S variable is t = 1 / (1 + p alpha r)
    float ek;       // -alpha^2 log2(e): exp(-(alpha r)^2) = 2^(ek r^2)
    float cg;       // 2 alpha / sqrt(pi)
    float sw_on;    // switch start, < 0 without a switch
    float sw_inv;   // 1 / (rc - r_switch)
    float sw_rc;    // the cutoff: 1 - t = (rc - r) sw_inv
    float sw_d;     // -30 / (rc - r_switch): dsv/dr = sw_d t^2 (1 - t)^2
};

__device__ __forceinline__ MixedPair mixed_pair_const(const SysDev& d) {
    MixedPair c;
    c.ap = (float)(kAsP * d.alpha);
    c.ek = (float)(-d.alpha * d.alpha * kLog2E);
    c.cg = (float)(kTwoOverSqrtPi * d.alpha);
    c.sw_on = (float)d.sw_on;
    c.sw_inv = (float)d.sw_inv;
    c.sw_rc = (float)d.rc;
    c.sw_d = (float)(-30.0 * d.sw_inv);
    return c;
}

// Mixed path: FP32 pair math.  Returns dE/dr / r (force = d * value on i).
// With erfc(alpha r) = ex u (ex = exp(-(alpha r)^2), u the A&S polynomial in
// t), the Coulomb part qq (cg ex + erfc / r) / r^2 is qq ex (cg + u / r) / r^2,
// so that it joins the LJ part in one multiply-add before the common 1/r^2.
__device__ __forceinline__ float pair_mixed(float r2, float qq, float sig, float epsij, const MixedPair& c,
                                            float& ecoul, float& elj) {
    const float inv_r = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * inv_r;
    const float inv_r2 = inv_r * inv_r;
    const float ex = __builtin_amdgcn_exp2f(c.ek * r2);                  // exp(-(alpha r)^2)
    const float u = erfc_as_t(__builtin_amdgcn_rcpf(fmaf(c.ap, r, 1.0f)), 1.0f);
    const float cy = ex * fmaf(inv_r, u, c.cg);
    const float s2 = sig * sig * inv_r2;
    const float s6 = s2 * s2 * s2;
    float le = epsij * (s6 - 1.0f) * s6;
    float dl = epsij * (12.0f * s6 - 6.0f) * s6;                          // r^2 times its dE/dr / r
    if (c.sw_on >= 0.f) {   // uniform per system
        // sv = 1 - 10 t^3 + 15 t^4 - 6 t^5 = w^3 (1 + 3 t + 6 t^2) with w = 1 - t,
        // dsv/dt = -30 t^2 w^2.  The factored form keeps FP32 accuracy where sv
        // is small (r near rc); w = 1 inside r_switch gives sv = 1, sd = 0.
        const float w = fminf((c.sw_rc - r) * c.sw_inv, 1.0f);
        const float t = 1.0f - w;
        const float sv = w * w * w * fmaf(t, fmaf(t, 6.0f, 3.0f), 1.0f);
        const float uw = t * w;
        const float sd = c.sw_d * uw * uw;
        dl = sv * dl - le * sd * r;
        le *= sv;
    }
    ecoul = qq * inv_r * ex * u;
    elj = le;
    return fmaf(qq, cy, dl) * inv_r2;
}

// FP64 path: the reference's formulas with library erfc/exp.
__device__ __forceinline__ double pair_fp64(double r2, double qq, double sig, double epsij, double alpha,
                                            double sw_on, double sw_inv, double& ecoul, double& elj) {
    const double r = sqrt(r2);
    const double inv_r = 1.0 / r;
    const double ar = alpha * r;
    const double erfcv = erfc(ar);
    const double ex = exp(-ar * ar);
    double dEdR = qq * inv_r * inv_r * inv_r * (erfcv + 2.0 * ar * ex / kSqrtPi);
    double s2 = sig * inv_r;
    s2 *= s2;
    const double s6 = s2 * s2 * s2;
    double le = epsij * (s6 - 1.0) * s6;
    double sv = 1.0, sd = 0.0;
    if (sw_on >= 0.0 && r > sw_on) {
        const double t = (r - sw_on) * sw_inv;
        sv = 1.0 + t * t * t * (-10.0 + t * (15.0 - t * 6.0));
        sd = t * t * (-30.0 + t * (60.0 - t * 30.0)) * sw_inv;
    }
    dEdR += sv * epsij * (12.0 * s6 - 6.0) * s6 * inv_r * inv_r;
    if (sw_on >= 0.0) {
        dEdR -= le * sd * inv_r;
        le *= sv;
    }
    ecoul = qq * inv_r * erfcv;
    elj = le;
    return dEdR;
}

// Wave-local ordering of LDS traffic (lanes of one wavefront exchanging data
// through LDS; the other wavefronts of the block run independently).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}


// Lanes l = 8 * ia + ja with ja <= ia: on a diagonal tile (i and j the same
// cluster) these are the self pairs and the half counted from the other side.
constexpr unsigned long long lower_triangle_mask() {
    unsigned long long m = 0ull;
    for (int l = 0; l < 64; ++l)
        if ((l & 7) <= (l >> 3)) m |= 1ull << l;
    return m;
}
constexpr unsigned long long kLowerTriangle = lower_triangle_mask();

template <class P, bool ENERGY>
__device__ __forceinline__ void pair_body(const SysDev* __restrict__ sd, const int* __restrict__ prefix,
                                                   int nsys, const int* __restrict__ wave_start,
                                                   const int4* __restrict__ chunks,
                                                   const uint2* __restrict__ list,
                                                   const unsigned long long* __restrict__ emask,
                                                   const typename P::real4* __restrict__ pos,
                                                   const typename P::real2* __restrict__ lj, typename P::facc* FX,
                                                   typename P::facc* FY, typename P::facc* FZ, long long* energy,
                                                   int* __restrict__ pool) {
    using real = typename P::real;
    using real4 = typename P::real4;
    using real2 = typename P::real2;
    // the wavefront index is the same on every lane: in a scalar register, the
    // chunk and list-entry indices derived from it are scalar too, and so are
    // the loads of chunk records and list words
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63, ia = lane >> 3,
              ja = lane & 7;
    __shared__ real4 ipos_s[kPairWaves][64];
    __shared__ real2 ilj_s[kPairWaves][64];
    __shared__ real jf_s[kPairWaves][kChunkEntries * 3 * 8];
    __shared__ uint2 jw_s[kPairWaves][kChunkEntries];
    real4* ipos = ipos_s[w];
    real2* ilj = ilj_s[w];
    real* jf = jf_s[w];
    uint2* jw = jw_s[w];

    // The list build cuts the batch's chunks into one contiguous range of equal
    // estimated work per wavefront.  The wavefronts of one XCD (blocks b with
    // b % 8 equal, as the hardware hands them out) pool their ranges and claim
    // chunks one at a time from an XCD counter until the pool is empty.  A fixed
    // range per wavefront finished 1.3x apart on one CU: the sequencer issues
    // for its oldest wavefront first, so the first block placed on a CU ran
    // ahead of the third.  Adjacent chunks are spatial neighbours, so an XCD's
    // L2 sees one region.  Every chunk is done whole by one wavefront: its
    // arithmetic does not depend on who does it or when.
    const int nb = gridDim.x, xcd = blockIdx.x & 7;
    const int per = (nb + 7) >> 3, full = (nb & 7) ? (nb & 7) : 8;     // blocks per XCD, as in xcd_remap
    const int xb = (xcd < full) ? xcd * per : full * per + (xcd - full) * (per - 1);
    const int xn = (xcd < full) ? per : per - 1;
    const int lo = wave_start[xb * kPairWaves], hi = wave_start[(xb + xn) * kPairWaves];
    int* counter = pool + xcd * kPoolStride;
    // lane 0 claims the next chunk; the result is read (and waited for) later
    auto claim = [&]() {
        int t = 0;
        if (lane == 0) t = atomicAdd(counter, 1);
        return t;
    };
    auto claimed = [&](int t) { return lo + __builtin_amdgcn_readfirstlane(t); };
    const real ke = (real)kOne4PiEps0;

    // Everything a chunk needs from memory before its first pair: the chunk
    // record, the i data of its clusters and its first list entry.  The next
    // chunk's copy is loaded before this chunk's force atomics are issued, so
    // the wait for it does not sit behind them (one in-order vmem counter).
    struct Pre {
        int s;
        int4 ch;
        real4 p;
        real2 l;
        uint2 words;     // lane l < kChunkEntries: the chunk's list word l
        uint2 ent;
        real4 pj;
        real2 ljj;
    };
    auto fetch = [&](int g) {
        Pre q;
        int lo = 0, hi = nsys - 1;     // largest s with prefix[s] <= g
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (prefix[mid] <= g) lo = mid; else hi = mid - 1;
        }
        q.s = lo;
        q.ch = chunks[sd[lo].chunk_off + (g - prefix[lo])];
        if ((lane >> 3) < q.ch.y) {
            q.p = pos[q.ch.x * 8 + lane];
            q.l = lj[q.ch.x * 8 + lane];
        } else {
            q.p.x = q.p.y = q.p.z = q.p.w = 0;
            q.l.x = q.l.y = 0;
        }
        // the first entry and its j data (none for a chunk without entries)
        q.words = make_uint2(0u, kNoMask);
        if (lane < q.ch.w - q.ch.z) q.words = list[q.ch.z + lane];
        q.ent = make_uint2(0u, kNoMask);
        q.pj.x = q.pj.y = q.pj.z = q.pj.w = 0;
        q.ljj.x = q.ljj.y = 0;
        if (q.ch.w > q.ch.z) {
            q.ent = list[q.ch.z];
            const int sj = (int)(q.ent.x & 0xffffffu) * 8 + ja;
            q.pj = pos[sj];
            q.ljj = lj[sj];
        }
        return q;
    };

    // The mixed path prefetches the next chunk across the flush below; the FP64
    // kernel is register-bound and fetches at the top of each chunk.
    int g = claimed(claim());
    Pre cur;
    if (P::mixed && g < hi) cur = fetch(g);
    while (g < hi) {
        if constexpr (!P::mixed) cur = fetch(g);
        // chunk values are the same on every lane: keep them in scalar registers
        const int s = __builtin_amdgcn_readfirstlane(cur.s);
        const int fc = __builtin_amdgcn_readfirstlane(cur.ch.x), ncl = __builtin_amdgcn_readfirstlane(cur.ch.y);
        const int e0 = __builtin_amdgcn_readfirstlane(cur.ch.z), e1 = __builtin_amdgcn_readfirstlane(cur.ch.w);
        // the claim of the chunk after this one: inside the entry loop, or here for
        // a chunk without entries (the loop below then does not run)
        int tk = 0;
        if (e1 <= e0) tk = claim();
        const real Lx = (real)sd[s].L[0], Ly = (real)sd[s].L[1], Lz = (real)sd[s].L[2];
        const real iLx = (real)sd[s].invL[0], iLy = (real)sd[s].invL[1], iLz = (real)sd[s].invL[2];
        const real rc2 = (real)sd[s].rc2;
        const real alpha = (real)sd[s].alpha;
        const real sw_on = (real)sd[s].sw_on, sw_inv = (real)sd[s].sw_inv;
        const MixedPair mpc = mixed_pair_const(sd[s]);   // mixed path only
        {
            real4 p = cur.p;
            p.w *= ke;
            ipos[lane] = p;
            ilj[lane] = cur.l;
            if (lane < kChunkEntries) jw[lane] = cur.words;
        }
        wave_lds_sync();
        real fxi[8], fyi[8], fzi[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) fxi[k] = fyi[k] = fzi[k] = 0;
        double ec_acc = 0.0, el_acc = 0.0;

        // software pipeline: the j data of entry e+1 loads while entry e computes.
        // The chunk's list words sit in LDS (written above with the i data); the
        // word of entry e+1 is read at the end of entry e-1.  A scalar load of it
        // would be simpler, but scalar loads return out of order, so every wait
        // for an LDS read issued after one would also wait for it.
        unsigned ex_n = __builtin_amdgcn_readfirstlane(cur.ent.x), ey_n = __builtin_amdgcn_readfirstlane(cur.ent.y);
        uint2 wv = jw[1];
        real4 pj = cur.pj;
        real2 lj_j = cur.ljj;
        for (int e = e0; e < e1; ++e) {
            // the list word is the same on every lane: in scalar registers, the
            // per-cluster tests below are scalar branches
            const unsigned ex = ex_n;
            const unsigned ey = ey_n;
            ex_n = __builtin_amdgcn_readfirstlane(wv.x);
            ey_n = __builtin_amdgcn_readfirstlane(wv.y);
            const int cj = (int)(ex & 0xffffffu);
            const unsigned im = ex >> 24;
            const unsigned eid = ey & kNoMask, img = ey >> kImgShift;
            // exclusion bits of the 8 cluster pairs, one load before the cluster loop
            unsigned long long em[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) em[k] = 0ull;
            if (eid != kNoMask) {
#pragma unroll
                for (int k = 0; k < 8; ++k) em[k] = emask[(size_t)eid * 8 + k];
            }
            // mixed path: the j cluster at the entry's periodic image (the list build
            // checked that it is the nearest image of every pair that can be within
            // the cutoff), or, for an entry whose pairs need it, the nearest image
            // pair by pair below; the FP64 path takes it pair by pair with the
            // reference's arithmetic, so that its cutoff test is the reference's bit
            // for bit.  Padding j slots move to -kPadPos, far from every i slot
            // (padding ones at +kPadPos included).
            const bool per_pair = !P::mixed || img == kImgPerPair;
            real4 pjs;
            {
                const real sx = per_pair ? (real)0 : ((img & 3u) == 1u ? Lx : ((img & 3u) == 2u ? -Lx : (real)0));
                const real sy = per_pair ? (real)0 : ((img & 12u) == 4u ? Ly : ((img & 12u) == 8u ? -Ly : (real)0));
                const real sz = per_pair ? (real)0 : ((img & 48u) == 16u ? Lz : ((img & 48u) == 32u ? -Lz : (real)0));
                pjs.x = (pj.x > (real)(0.5 * kPadPos)) ? (real)(-kPadPos) : pj.x + sx;
                pjs.y = pj.y + sy;
                pjs.z = pj.z + sz;
                pjs.w = pj.w;
            }
            real4 pj_n = pj;
            real2 lj_n = lj_j;
            if (e + 1 < e1) {
                const int sjn = (int)(ex_n & 0xffffffu) * 8 + ja;
                pj_n = pos[sjn];
                lj_n = lj[sjn];
            }
            // claim the next chunk two entries before this chunk's end, after the
            // loads above: the waits for them then do not wait for the atomic (the
            // vector memory counter drains in issue order), and its round trip
            // overlaps the last entries' pairs
            if (e == max(e0, e1 - 2)) tk = claim();
            real fxj = 0, fyj = 0, fzj = 0;
            // i data of each cluster is read from LDS at its tile (reading it one
            // tile ahead cost 6 registers and measured slower; the other waves on
            // the SIMD hide the read).  The empty asm makes the LDS index opaque per
            // entry: otherwise the compiler hoists all 8 clusters' reads out of the
            // entry loop and holds them in 48 more registers, which spills.
            int io = ia;
            asm volatile("" : "+v"(io));
            const real4* ipe = ipos + io;
            const real2* ile = ilj + io;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (im & (1u << k)) {
                    const real4 pi = ipe[k * 8];
                    const real2 li = ile[k * 8];
                    // excluded lanes, as a wave-wide bit mask: listed exclusions, and on
                    // the diagonal tile the pairs with ja <= ia (self pairs and the other half)
                    const unsigned long long xm = em[k] | (cj == fc + k ? kLowerTriangle : 0ull);
                    real dx = pi.x - pjs.x, dy = pi.y - pjs.y, dz = pi.z - pjs.z;
                    if (per_pair) {     // a uniform branch
                        if constexpr (P::mixed) {
                            // nearest image by rounding (equal to the one-correction rule for
                            // |d| < 1.5 L); a padding slot keeps its distance of ~kPadPos in x
                            const real cx = dx - Lx * rintf(dx * iLx);
                            dx = fabsf(dx) < (real)(0.5 * kPadPos) ? cx : dx;
                            dy -= Ly * rintf(dy * iLy);
                            dz -= Lz * rintf(dz * iLz);
                        } else {
                            dx = min_image1(dx, Lx);
                            dy = min_image1(dy, Ly);
                            dz = min_image1(dz, Lz);
                        }
                    }
                    real r2;
                    if constexpr (P::mixed) r2 = dx * dx + dy * dy + dz * dz;
                    else r2 = dot_rn(dx, dy, dz);
                    // lane l is on when bit l of the cutoff ballot is set and bit l of the
                    // exclusion mask is clear: the mask logic runs in scalar registers
                    const unsigned long long onm = __builtin_amdgcn_ballot_w64(r2 < rc2) & ~xm;
                    // a tile with no pair inside the cutoff (kept by the list's outer
                    // radius) skips the pair arithmetic: a uniform branch
                    if constexpr (P::mixed) {
                        if (onm == 0ull) continue;
                    }
                    const bool on = __builtin_amdgcn_inverse_ballot_w64(onm);
                    real ecv, elv, dEdR;
                    if constexpr (P::mixed) {
                        // evaluated on every lane; lanes that are off (incl. r2 = 0) select zero
                        dEdR = pair_mixed(r2, pi.w * pjs.w, li.x + lj_j.x, li.y * lj_j.y, mpc, ecv, elv);
                        dEdR = on ? dEdR : 0.0f;
                    } else {
                        dEdR = 0;
                        ecv = elv = 0;
                        if (on)
                            dEdR = pair_fp64(r2, pi.w * pjs.w, li.x + lj_j.x, li.y * lj_j.y, alpha, sw_on,
                                             sw_inv, ecv, elv);
                    }
                    // every displacement is finite (padding slots included), so the
                    // zero of an off lane gives zero force
                    const real fx = dx * dEdR, fy = dy * dEdR, fz = dz * dEdR;
                    fxi[k] += fx; fyi[k] += fy; fzi[k] += fz;
                    fxj -= fx; fyj -= fy; fzj -= fz;
                    if constexpr (ENERGY) {
                        if (on) { ec_acc += (double)ecv; el_acc += (double)elv; }
                    }
                }
            }
            fxj = sum_stride8(fxj);
            fyj = sum_stride8(fyj);
            fzj = sum_stride8(fzj);
            // staged in LDS: no global atomics inside the loop, so waiting for the
            // prefetched loads never waits behind atomics
            const int el = e - e0;
            if (ia == 0) {
                jf[(el * 3 + 0) * 8 + ja] = fxj;
                jf[(el * 3 + 1) * 8 + ja] = fyj;
                jf[(el * 3 + 2) * 8 + ja] = fzj;
            }
            if (e + 2 < e1) wv = jw[el + 2];
            pj = pj_n;
            lj_j = lj_n;
        }
        const int gn = claimed(tk);
        if constexpr (P::mixed) {
            if (gn < hi) cur = fetch(gn);
        }
        wave_lds_sync();
        {
            // j-forces of the chunk, 8 entries (64 lanes) per atomic instruction
            const int n = e1 - e0;
            for (int b = 0; b < n; b += 8) {
                const int el = b + ia;
                if (el < n) {
                    const int sj = (int)(jw[el].x & 0xffffffu) * 8 + ja;
                    facc_add(&FX[sj], jf[(el * 3 + 0) * 8 + ja]);
                    facc_add(&FY[sj], jf[(el * 3 + 1) * 8 + ja]);
                    facc_add(&FZ[sj], jf[(el * 3 + 2) * 8 + ja]);
                }
            }
        }
        // i-forces: sums over the 8 j lanes, transposed through LDS so that one
        // full-wave atomic per component covers the chunk's 64 i atoms
        wave_lds_sync();   // the j flush has read jf
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const real fx = sum_group8(fxi[k]), fy = sum_group8(fyi[k]), fz = sum_group8(fzi[k]);
            if (ja == 0) {
                jf[0 * 64 + k * 8 + ia] = fx;
                jf[1 * 64 + k * 8 + ia] = fy;
                jf[2 * 64 + k * 8 + ia] = fz;
            }
        }
        wave_lds_sync();
        if ((lane >> 3) < ncl) {
            const int slot = fc * 8 + lane;
            facc_add(&FX[slot], jf[0 * 64 + lane]);
            facc_add(&FY[slot], jf[1 * 64 + lane]);
            facc_add(&FZ[slot], jf[2 * 64 + lane]);
        }
        if constexpr (ENERGY) {
            const long long ec = wave_sum(energy_fixed(ec_acc));
            const long long el = wave_sum(energy_fixed(el_acc));
            if (lane == 0) {
                energy_add_fixed(&energy[s * E_NTERMS + E_COUL], ec);
                energy_add_fixed(&energy[s * E_NTERMS + E_LJ], el);
            }
        }
        wave_lds_sync();   // LDS of this wave is reused by its next chunk
        g = gn;
    }
}

#define NX_PAIR_PARAMS(P)                                                                                      \
    const SysDev *__restrict__ sd, const int *__restrict__ prefix, int nsys, const int *__restrict__ wave_start,  \
        const int4 *__restrict__ chunks, const uint2 *__restrict__ list, const unsigned long long *__restrict__ emask, \
        const typename P::real4 *__restrict__ pos, const typename P::real2 *__restrict__ lj, typename P::facc *FX,    \
        typename P::facc *FY, typename P::facc *FZ, long long *energy, int *__restrict__ pool
#define NX_PAIR_ARGS sd, prefix, nsys, wave_start, chunks, list, emask, pos, lj, FX, FY, FZ, energy, pool

// The pair kernels.  The mixed one without energies must stay within
// PrecMixed::pair_max_vgpr VGPRs (the compiler offers no cap short of a higher
// occupancy target, which spills; a test of the build checks the count), so
// that the PME kernels still find room beside its persistent grid.
template <bool ENERGY>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PrecMixed::pair_waves_per_simd))) void
pair_kernel_mixed(NX_PAIR_PARAMS(PrecMixed)) {
    pair_body<PrecMixed, ENERGY>(NX_PAIR_ARGS);
}
template <bool ENERGY>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PrecFP64::pair_waves_per_simd))) void
pair_kernel_fp64(NX_PAIR_PARAMS(PrecFP64)) {
    pair_body<PrecFP64, ENERGY>(NX_PAIR_ARGS);
}
#undef NX_PAIR_ARGS
#undef NX_PAIR_PARAMS

template <class P, bool ENERGY>
constexpr auto pair_kernel_of() {
    if constexpr (P::mixed) return pair_kernel_mixed<ENERGY>;
    else return pair_kernel_fp64<ENERGY>;
}

// Explicit pair set (atom ids, first > second) with the FP64 test of the pair kernel.
template <class P>
__global__ __launch_bounds__(64) void export_pairs_kernel(const SysDev* sd, const SysCount* sc, const int* super_sys,
                                                          const int4* super_item, const uint2* list,
                                                          const unsigned long long* emask, const int* atom_of_slot,
                                                          const double* X, const double* Y, const double* Z,
                                                          int2* out, long long cap, unsigned long long* count,
                                                          int n_items) {
    const int item = blockIdx.x;
    if (item >= n_items) return;
    const int s = super_sys[item];
    if (item - sd[s].super_off >= sc[s].n_super) return;
    const int4 it = super_item[item];
    const int fc = it.x, ncl = it.y, lstart = it.z, llen = it.w;
    const int lane = threadIdx.x, ia = lane >> 3, ja = lane & 7;
    const double Lx = sd[s].L[0], Ly = sd[s].L[1], Lz = sd[s].L[2], rc2 = sd[s].rc2;
    for (int e = 0; e < llen; ++e) {
        const uint2 ent = list[lstart + e];
        const int cj = (int)(ent.x & 0xffffffu);
        const unsigned im = ent.x >> 24;
        const int aj = atom_of_slot[cj * 8 + ja];
        for (int k = 0; k < ncl; ++k) {
            if (!(im & (1u << k))) continue;
            const int ai = atom_of_slot[(fc + k) * 8 + ia];
            bool excl = (cj == fc + k) && (ja <= ia);
            const unsigned eid = ent.y & kNoMask;
            if (eid != kNoMask) excl |= ((emask[(size_t)eid * 8 + k] >> lane) & 1ull) != 0ull;
            if (excl || ai < 0 || aj < 0) continue;
            const double dx = min_image1(X[ai] - X[aj], Lx);
            const double dy = min_image1(Y[ai] - Y[aj], Ly);
            const double dz = min_image1(Z[ai] - Z[aj], Lz);
            if (dot_rn(dx, dy, dz) < rc2) {
                unsigned long long w = atomicAdd(count, 1ull);
                if ((long long)w < cap) out[w] = make_int2(max(ai, aj), min(ai, aj));
            }
        }
    }
}

// Blocks of the pair kernel one CU holds at once (its occupancy limit).
template <class P>
static int pair_blocks_per_cu() {
    static const int v = [] {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, pair_kernel_of<P, false>(), 256, 0) != hipSuccess || n <= 0)
            n = 2;
        return n;
    }();
    return v;
}

// Persistent grid of the pair kernel, all blocks resident at once, at most
// three per CU: the rest of a CU's wave slots go to the bonded and PME kernels
// running beside it.  Measured on the 100k-atom step (mixed, five resident
// blocks possible): 3 per CU 261 us, 4 per CU 303 us, 2 per CU 268 us.
// Fixed per batch layout: the list build cuts the chunk ranges for it.
template <class P>
int pair_grid_blocks(int n_cu, int n_chunk_cap, int per_cu_request) {
    // a requested count (the autotuner's) is capped by what a CU can hold at once
    const int occ = pair_blocks_per_cu<P>();
    const int per_cu = per_cu_request > 0 ? std::min(per_cu_request, occ) : std::min(occ, 3);
    return std::max(1, std::min((n_chunk_cap + kPairWaves - 1) / kPairWaves, per_cu * std::max(1, n_cu)));
}

// Ends the step's list work before the pair kernel: clears the rebuild flags
// (every list kernel that reads them ran before) and counts the builds.  As a
// separate launch it also gives the kernels forked to the side stream (bonded
// terms, PME) a head start on the CUs before the pair kernel's persistent grid
// takes its slots; without it the step measured about 10% slower.
__global__ void pair_begin_kernel(SysCount* sc, int nsys, int* pool) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < 8) pool[s * kPoolStride] = 0;     // no chunk of the pair kernel claimed yet
    if (s >= nsys) return;
    if (sc[s].rebuild) {
        sc[s].rebuild = 0;
        sc[s].rebuilds += 1;
    }
}

template <class P>
void launch_pair_begin(Batch<P>& B) {
    pair_begin_kernel<<<(B.nsys + 63) / 64, 64, 0, B.stream>>>(B.d_cnt, B.nsys, B.pair_pool);
}

template <class P>
void launch_pair_kernel(Batch<P>& B, bool energy) {
    if (B.n_chunk_cap == 0) return;
    const unsigned blocks = (unsigned)B.pair_blocks;
    if (energy)
        pair_kernel_of<P, true>()<<<blocks, 256, 0, B.stream>>>(B.d_sys, B.chunk_prefix, B.nsys, B.wave_start, B.chunks, B.list,
                                                           B.emask, B.pos, B.lj, B.FSx, B.FSy, B.FSz, B.energy,
                                                           B.pair_pool);
    else
        pair_kernel_of<P, false>()<<<blocks, 256, 0, B.stream>>>(B.d_sys, B.chunk_prefix, B.nsys, B.wave_start, B.chunks, B.list,
                                                            B.emask, B.pos, B.lj, B.FSx, B.FSy, B.FSz, B.energy,
                                                            B.pair_pool);
}

template <class P>
void launch_pair(Batch<P>& B, bool energy) {
    launch_pair_begin<P>(B);
    launch_pair_kernel<P>(B, energy);
}

template <class P>
void launch_export_pairs(Batch<P>& B, int2* out, long long cap, long long* d_count) {
    export_pairs_kernel<P><<<B.n_super_cap, 64, 0, B.stream>>>(B.d_sys, B.d_cnt, B.super_sys, B.super_item, B.list,
                                                               B.emask, B.atom_of_slot, B.X, B.Y, B.Z, out, cap,
                                                               reinterpret_cast<unsigned long long*>(d_count),
                                                               B.n_super_cap);
}

template int pair_grid_blocks<PrecFP64>(int, int, int);
template int pair_grid_blocks<PrecMixed>(int, int, int);

template <class P>
int pair_blocks_occupancy() {
    return pair_blocks_per_cu<P>();
}
template int pair_blocks_occupancy<PrecFP64>();
template int pair_blocks_occupancy<PrecMixed>();
template void launch_pair<PrecFP64>(Batch<PrecFP64>&, bool);
template void launch_pair<PrecMixed>(Batch<PrecMixed>&, bool);
template void launch_pair_begin<PrecFP64>(Batch<PrecFP64>&);
template void launch_pair_begin<PrecMixed>(Batch<PrecMixed>&);
template void launch_pair_kernel<PrecFP64>(Batch<PrecFP64>&, bool);
template void launch_pair_kernel<PrecMixed>(Batch<PrecMixed>&, bool);
template void launch_export_pairs<PrecFP64>(Batch<PrecFP64>&, int2*, long long, long long*);
template void launch_export_pairs<PrecMixed>(Batch<PrecMixed>&, int2*, long long, long long*);

}  // namespace nx

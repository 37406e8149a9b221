/*
 * nxo.c -- CPU FP64 oracle (test infrastructure only; see nxo.h).
 *
 * Plain-C restatement of NgFEP/NexusMD's single-threaded engine.  Operation
 * order follows the reference where it matters for rounding (dot products
 * left to right, no FMA contraction: build with -ffp-contract=off).
 * The 3D FFT restates the discrete Fourier transform that the reference
 * obtains from FFTW 3.3.5 (core/libs/fftw-3.3.5-dll64, called at
 * core/baseline/src/Forces/NonbondedForce.cpp:264-290): in-place C2C,
 * unnormalised, sign -1 forward and +1 backward.  FFTW itself is absent
 * (Win64 binary only), so this is a mixed-radix Cooley-Tukey written from
 * the DFT definition.
 */
#include "nxo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

void nxo_free(void* p) { free(p); }

/* ------------------------------------------------------------------ box */

/* core/common/src/PeriodicBoundaryCondition.cpp:11-88 (water_availibility
 * is hard-wired true there, so only the 1% margin branch :75-86 runs). */
void nxo_box_from_positions(int n, const double* pos, nxo_box* box) {
    double lb[3], ub[3], sz[3];
    for (int d = 0; d < 3; ++d) lb[d] = ub[d] = pos[d];
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            double v = pos[3 * i + d];
            if (v < lb[d]) lb[d] = v;      /* std::min(lb, v) */
            if (ub[d] < v) ub[d] = v;      /* std::max(ub, v) */
        }
    for (int d = 0; d < 3; ++d) sz[d] = fabs(ub[d] - lb[d]);
    for (int d = 0; d < 3; ++d) {
        double margin = 0.01 * sz[d];
        lb[d] -= margin;
        ub[d] += margin;
    }
    nxo_box_from_bounds(lb, ub, box);
}

void nxo_box_from_bounds(const double lb[3], const double ub[3], nxo_box* box) {
    for (int d = 0; d < 3; ++d) {
        box->lb[d] = lb[d];
        box->ub[d] = ub[d];
        box->L[d] = fabs(ub[d] - lb[d]);
    }
}

/* PeriodicBoundaryCondition.cpp:93-104: d = a - b, one +-L correction. */
void nxo_min_image(const double* a, const double* b, const nxo_box* box, double* d) {
    for (int k = 0; k < 3; ++k) {
        double v = a[k] - b[k];
        if (v > 0.5 * box->L[k]) v -= box->L[k];
        else if (v < -0.5 * box->L[k]) v += box->L[k];
        d[k] = v;
    }
}

static inline double dot3(const double* a, const double* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];   /* Coords3D.h:116-124 */
}
static inline void cross3(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];                 /* Coords3D.h:126-128 */
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* ----------------------------------------------------------- exclusions */

typedef struct { int n, cap; int* v; } iset;   /* sorted int set */

static void iset_insert(iset* s, int x) {
    int lo = 0, hi = s->n;
    while (lo < hi) { int m = (lo + hi) / 2; if (s->v[m] < x) lo = m + 1; else hi = m; }
    if (lo < s->n && s->v[lo] == x) return;
    if (s->n == s->cap) { s->cap = s->cap ? 2 * s->cap : 8; s->v = (int*)realloc(s->v, sizeof(int) * s->cap); }
    memmove(s->v + lo + 1, s->v + lo, sizeof(int) * (s->n - lo));
    s->v[lo] = x;
    s->n++;
}

/* core/baseline/src/Forces/Exclusions.cpp:7-44 (bondCutoff = 3 from
 * core/baseline/headers/Engine.h:58).  Returns -1 on an illegal index
 * (the reference throws runtime_error, Exclusions.cpp:12-15). */
int nxo_exclusions(int n, int nb, const nxo_bond* bonds, int bond_cutoff, int** off_out, int** idx_out) {
    for (int b = 0; b < nb; ++b)
        if (bonds[b].p1 < 0 || bonds[b].p2 < 0 || bonds[b].p1 >= n || bonds[b].p2 >= n) return -1;
    iset* ex = (iset*)calloc(n, sizeof(iset));
    iset* b12 = (iset*)calloc(n, sizeof(iset));
    if (bond_cutoff >= 1) {
        for (int b = 0; b < nb; ++b) {
            int p1 = bonds[b].p1, p2 = bonds[b].p2;
            iset_insert(&ex[p1], p2); iset_insert(&ex[p2], p1);
            iset_insert(&b12[p1], p2); iset_insert(&b12[p2], p1);
        }
        for (int level = 0; level < bond_cutoff - 1; ++level) {
            /* currentExclusions = exclusions (a full copy per level) */
            iset* cur = (iset*)calloc(n, sizeof(iset));
            for (int i = 0; i < n; ++i) {
                cur[i].n = cur[i].cap = ex[i].n;
                cur[i].v = (int*)malloc(sizeof(int) * (ex[i].n ? ex[i].n : 1));
                memcpy(cur[i].v, ex[i].v, sizeof(int) * ex[i].n);
            }
            for (int i = 0; i < n; ++i)
                for (int a = 0; a < cur[i].n; ++a) {
                    int j = cur[i].v[a];
                    for (int c = 0; c < b12[i].n; ++c) {
                        int bonded = b12[i].v[c];
                        if (bonded != j) iset_insert(&ex[j], bonded);
                    }
                }
            for (int i = 0; i < n; ++i) free(cur[i].v);
            free(cur);
        }
    }
    int* off = (int*)malloc(sizeof(int) * (n + 1));
    off[0] = 0;
    for (int i = 0; i < n; ++i) off[i + 1] = off[i] + ex[i].n;
    int* idx = (int*)malloc(sizeof(int) * (off[n] ? off[n] : 1));
    for (int i = 0; i < n; ++i) { memcpy(idx + off[i], ex[i].v, sizeof(int) * ex[i].n); free(ex[i].v); free(b12[i].v); }
    free(ex); free(b12);
    *off_out = off; *idx_out = idx;
    return off[n];
}

static int is_excluded(const int* off, const int* idx, int i, int j) {
    if (!off) return 0;
    int lo = off[i], hi = off[i + 1];
    while (lo < hi) { int m = (lo + hi) / 2; if (idx[m] < j) lo = m + 1; else hi = m; }
    return lo < off[i + 1] && idx[lo] == j;
}

/* -------------------------------------------------------- neighbour list */

/* Restatement of core/baseline/src/Forces/NeighborList.cpp:12-154
 * (VoxelHash + obtainNeighborList, usePeriodic = true, minDistance = 0).
 * The std::map of voxels becomes a dense table over voxel keys in
 * [-1, n] per axis; a lookup of a key the reference would not find (a
 * wrapped index that is negative, or the unwrapped key n that insert() can
 * produce for r within an ulp of L) finds nothing here either. */
int64_t nxo_pairs_voxel(int n, const double* pos, const nxo_box* box, const int* excl_off,
                        const int* excl_idx, double cutoff, int* pairs, int64_t cap) {
    const double* L = box->L;
    double vs[3];
    int nv[3];
    for (int d = 0; d < 3; ++d) {
        double edge = 0.5 * L[d] / floor(L[d] / cutoff);        /* :146-148 */
        nv[d] = (int)floor(L[d] / edge + 0.5);                   /* :15-17 */
        vs[d] = L[d] / nv[d];                                    /* :18-20 */
    }
    const int W0 = nv[0] + 2, W1 = nv[1] + 2, W2 = nv[2] + 2;   /* keys -1..n */
    int64_t nvox = (int64_t)W0 * W1 * W2;
    int* head = (int*)malloc(sizeof(int) * nvox);
    int* tail = (int*)malloc(sizeof(int) * nvox);
    for (int64_t v = 0; v < nvox; ++v) head[v] = tail[v] = -1;
    int* next = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
    int* key = (int*)malloc(sizeof(int) * 3 * (n > 0 ? n : 1));
    const double max2 = cutoff * cutoff;
    int dI[3];
    for (int d = 0; d < 3; ++d) dI[d] = (int)(cutoff / vs[d]) + 1;   /* :58-60 */
    int64_t np = 0;
    for (int a = 0; a < n; ++a) {
        const double* li = pos + 3 * a;
        int c[3];
        for (int d = 0; d < 3; ++d) {                            /* getVoxelIndex :24-35 */
            double r = li[d];
            r -= floor(r / L[d]) * L[d];
            c[d] = (int)floor(r / vs[d]);
        }
        int minz = c[2] - dI[2], maxz = c[2] + dI[2];
        if (maxz > minz + nv[2] - 1) maxz = minz + nv[2] - 1;
        for (int z = minz; z <= maxz; ++z) {
            int miny = c[1] - dI[1], maxy = c[1] + dI[1];
            if (maxy > miny + nv[1] - 1) maxy = miny + nv[1] - 1;
            for (int y = miny; y <= maxy; ++y) {
                int minx = c[0] - dI[0], maxx = c[0] + dI[0];
                if (maxx > minx + nv[0] - 1) maxx = minx + nv[0] - 1;
                for (int x = minx; x <= maxx; ++x) {
                    int kx = (x + nv[0]) % nv[0], ky = (y + nv[1]) % nv[1], kz = (z + nv[2]) % nv[2];
                    if (kx < -1 || ky < -1 || kz < -1) continue;       /* not a stored key */
                    int64_t v = ((int64_t)(kx + 1) * W1 + (ky + 1)) * W2 + (kz + 1);
                    for (int b = head[v]; b >= 0; b = next[b]) {
                        double dr[3];
                        nxo_min_image(pos + 3 * b, li, box, dr);       /* J - I, :106 */
                        double r2 = dot3(dr, dr);
                        if (is_excluded(excl_off, excl_idx, a, b)) continue;
                        if (r2 >= 0.0 && r2 < max2) {
                            if (np < cap) { pairs[2 * np] = a; pairs[2 * np + 1] = b; }
                            np++;
                        }
                    }
                }
            }
        }
        /* insert after the query (:151-152), at the end of the voxel's list: a
           Voxel is a vector filled by push_back (:37-42), visited in insertion order */
        int k0 = c[0] < -1 ? -1 : c[0], k1 = c[1] < -1 ? -1 : c[1], k2 = c[2] < -1 ? -1 : c[2];
        if (k0 > nv[0]) k0 = nv[0];
        if (k1 > nv[1]) k1 = nv[1];
        if (k2 > nv[2]) k2 = nv[2];
        key[3 * a] = k0; key[3 * a + 1] = k1; key[3 * a + 2] = k2;
        int64_t v = ((int64_t)(k0 + 1) * W1 + (k1 + 1)) * W2 + (k2 + 1);
        next[a] = -1;
        if (tail[v] >= 0) next[tail[v]] = a;
        else head[v] = a;
        tail[v] = a;
    }
    free(head); free(tail); free(next); free(key);
    return np;
}

/* The mathematical pair set: every non-excluded (i > j) with the reference's
 * single-correction minimum-image r^2 < rc^2 (J - I, left-to-right dot). */
int64_t nxo_pairs_brute(int n, const double* pos, const nxo_box* box, const int* excl_off,
                        const int* excl_idx, double cutoff, int* pairs, int64_t cap) {
    const double max2 = cutoff * cutoff;
    int64_t np = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) {
            double dr[3];
            nxo_min_image(pos + 3 * j, pos + 3 * i, box, dr);
            double r2 = dot3(dr, dr);
            if (r2 < max2 && !is_excluded(excl_off, excl_idx, i, j)) {
                if (np < cap) { pairs[2 * np] = i; pairs[2 * np + 1] = j; }
                np++;
            }
        }
    return np;
}

/* -------------------------------------------------------------- bonded */

/* core/baseline/src/Forces/HarmonicBondForce.cpp:20-58 via Forces.cpp:42-52 */
double nxo_bonds(int nb, const nxo_bond* b, const double* pos, const nxo_box* box, double* f) {
    double e = 0.0;
    for (int t = 0; t < nb; ++t) {
        const double* p1 = pos + 3 * b[t].p1;
        const double* p2 = pos + 3 * b[t].p2;
        double d[3];
        nxo_min_image(p2, p1, box, d);
        double r = sqrt(dot3(d, d));
        double dI = r - b[t].d;
        e += 0.5 * b[t].k * dI * dI;
        double dEdR = b[t].k * dI;
        if (r > 0) { double s = dEdR / r; d[0] *= s; d[1] *= s; d[2] *= s; }
        else d[0] = d[1] = d[2] = 0.0;
        for (int k = 0; k < 3; ++k) { f[3 * b[t].p1 + k] += d[k]; f[3 * b[t].p2 + k] += -d[k]; }
    }
    return e;
}

/* core/baseline/src/Forces/HarmonicAngleForce.cpp:19-63 via Forces.cpp:54-66 */
double nxo_angles(int na, const nxo_angle* a, const double* pos, const nxo_box* box, double* f) {
    double e = 0.0;
    for (int t = 0; t < na; ++t) {
        const double* p1 = pos + 3 * a[t].p1;
        const double* p2 = pos + 3 * a[t].p2;
        const double* p3 = pos + 3 * a[t].p3;
        double v0[3], v1[3], cp[3];
        nxo_min_image(p2, p1, box, v0);
        nxo_min_image(p2, p3, box, v1);
        /* calculateAngle (:19-35) */
        double r21 = sqrt(dot3(v0, v0));
        double r23 = sqrt(dot3(v1, v1));
        double dt = dot3(v0, v1);
        double c = dt / (r21 * r23);
        if (c < -1.0) c = -1.0;
        if (c > 1.0) c = 1.0;
        double theta = acos(c);
        double dI = theta - a[t].a;
        e += 0.5 * a[t].k * dI * dI;
        double dE = a[t].k * dI;
        /* calculateForces (:46-59) */
        cross3(v0, v1, cp);
        double rp = sqrt(dot3(cp, cp));
        if (rp < 1.0e-06) rp = 1.0e-06;
        double q21 = dot3(v0, v0), q23 = dot3(v1, v1);
        double f1[3], f3[3];
        cross3(v0, cp, f1);
        cross3(cp, v1, f3);
        double s1 = dE / (q21 * rp), s3 = dE / (q23 * rp);
        for (int k = 0; k < 3; ++k) {
            f1[k] *= s1; f3[k] *= s3;
            double f2 = -f1[k] - f3[k];
            f[3 * a[t].p1 + k] += f1[k];
            f[3 * a[t].p2 + k] += f2;
            f[3 * a[t].p3 + k] += f3[k];
        }
    }
    return e;
}

/* Coords3D::normalize (Coords3D.h:131-141): divides by the length. */
static void normalize3(const double* v, double* o) {
    double len = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    o[0] = o[1] = o[2] = 0.0;
    if (len > 0) { o[0] = v[0] / len; o[1] = v[1] / len; o[2] = v[2] / len; }
}

/* core/baseline/src/Forces/PeriodicTorsionForce.cpp:41-149 via Forces.cpp:18-39 */
double nxo_torsions(int nt, const nxo_torsion* t, const double* pos, const nxo_box* box, double* f) {
    const double thr = (double)0.99f;   /* `cosangle > 0.99f` compares against the float literal */
    double e = 0.0;
    for (int u = 0; u < nt; ++u) {
        const double* r1 = pos + 3 * t[u].p1;
        const double* r2 = pos + 3 * t[u].p2;
        const double* r3 = pos + 3 * t[u].p3;
        const double* r4 = pos + 3 * t[u].p4;
        double v0[3], v1[3], v2[3], cp0[3], cp1[3];
        nxo_min_image(r1, r2, box, v0);
        nxo_min_image(r3, r2, box, v1);
        nxo_min_image(r3, r4, box, v2);
        cross3(v0, v1, cp0);
        cross3(v1, v2, cp1);
        /* torsion_angle (:41-77) */
        double n0[3], n1[3];
        normalize3(cp0, n0);
        normalize3(cp1, n1);
        double cosang = dot3(n0, n1);
        double ang;
        if (cosang > thr || cosang < -thr) {
            double cr[3];
            cross3(cp0, cp1, cr);
            double scale = dot3(cp0, cp0) * dot3(cp1, cp1);
            ang = asin(sqrt(dot3(cr, cr) / scale));
            if (cosang < 0) ang = NXO_PI - ang;
        } else {
            ang = acos(cosang);
        }
        ang = (dot3(v0, cp1) >= 0 ? ang : -ang);
        /* calculateForces (:81-149) */
        double delta = t[u].periodicity * ang - t[u].phase;
        e += t[u].k * (1.0 + cos(delta));
        double dEdPhi = -t[u].k * t[u].periodicity * sin(delta);
        double nc1 = dot3(cp0, cp0);
        double nsq = dot3(v1, v1);
        double nbc = sqrt(nsq);
        double nc2 = dot3(cp1, cp1);
        double dp = 1.0 / nsq;
        double fa1 = (-dEdPhi * nbc) / nc1;
        double fa2 = dot3(v0, v1) * dp;
        double fa3 = dot3(v2, v1) * dp;
        double fa4 = (dEdPhi * nbc) / nc2;
        for (int k = 0; k < 3; ++k) {
            double F1 = cp0[k] * fa1;
            double F4 = cp1[k] * fa4;
            double s = F1 * fa2 - F4 * fa3;
            f[3 * t[u].p1 + k] += F1;
            f[3 * t[u].p2 + k] += s - F1;
            f[3 * t[u].p3 + k] += -s - F4;
            f[3 * t[u].p4 + k] += F4;
        }
    }
    return e;
}

/* ---------------------------------------------------------------- PME */

/* NonbondedForce.cpp:293-321, "approximation" branch (:297-300). */
double nxo_alpha(double cutoff, double tol) { return (1.0 / cutoff) * sqrt(-log(2.0 * tol)); }

/* NonbondedForce.cpp:33-48: next power of two >= 10*ceil(L). */
void nxo_grid_rule(const nxo_box* box, int grid[3]) {
    for (int d = 0; d < 3; ++d) {
        int m = (int)ceil(box->L[d]) * 10;
        int p = 1;
        while (p < m) p *= 2;
        grid[d] = p;
    }
}

/* B-spline values at integer knots, normalised (NonbondedForce.cpp:77-110). */
static void bspline_knots(int order, double* data) {
    for (int i = 0; i < order; ++i) data[i] = 0.0;
    data[order - 1] = 0.0;
    data[1] = 0.0;
    data[0] = 1.0;
    for (int i = 3; i < order; ++i) {
        double div = 1.0 / (i - 1);
        data[i - 1] = 0.0;
        for (int j = 1; j < (i - 1); ++j)
            data[i - j - 1] = div * (j * data[i - j - 2] + (i - j) * data[i - j - 1]);
        data[0] = div * data[0];
    }
    double div = 1.0 / (order - 1);
    data[order - 1] = div * data[order - 2];
    for (int i = 1; i < (order - 1); ++i)
        data[order - i - 1] = div * (i * data[order - i - 2] + (order - i) * data[order - i - 1]);
    data[0] = div * data[0];
}

/* |b(m)|^2 moduli by an O(K^2) DFT with the neighbour-average fix-up
 * (NonbondedForce.cpp:112-142; for order 5 on even K this fires at K/2). */
void nxo_bspline_moduli(int K, int order, double* mod) {
    double data[16];
    bspline_knots(order, data);
    int M = K > order + 1 ? K : order + 1;
    double* bd = (double*)calloc(M, sizeof(double));
    for (int i = 1; i <= order; ++i) bd[i] = data[i - 1];
    for (int i = 0; i < K; ++i) {
        double sc = 0.0, ss = 0.0;
        for (int j = 0; j < K; ++j) {
            double arg = (2 * NXO_PI * i * j) / K;
            sc += bd[j] * cos(arg);
            ss += bd[j] * sin(arg);
        }
        mod[i] = sc * sc + ss * ss;
    }
    for (int i = 0; i < K; ++i)
        if (mod[i] < 1e-7) mod[i] = (mod[(i - 1 + K) % K] + mod[(i + 1) % K]) / 2.0;
    free(bd);
}

/* theta / dtheta for one fractional offset (NonbondedForce.cpp:150-181). */
void nxo_bspline_theta(double dr, int order, double* th, double* dth) {
    th[order - 1] = 0;
    th[1] = dr;
    th[0] = 1 - dr;
    for (int k = 3; k < order; ++k) {
        double div = 1.0 / (k - 1.0);
        th[k - 1] = div * dr * th[k - 2];
        for (int l = 1; l < (k - 1); ++l)
            th[k - l - 1] = div * ((dr + l) * th[k - l - 2] + (k - l - dr) * th[k - l - 1]);
        th[0] = div * (1 - dr) * th[0];
    }
    dth[0] = -th[0];
    for (int k = 1; k < order; ++k) dth[k] = th[k - 1] - th[k];
    double div = 1.0 / (order - 1);
    th[order - 1] = div * dr * th[order - 2];
    for (int l = 1; l < (order - 1); ++l)
        th[order - l - 1] = div * ((dr + l) * th[order - l - 2] + (order - l - dr) * th[order - l - 1]);
    th[0] = div * (1 - dr) * th[0];
}

/* NonbondedForce.cpp:51-71 (origin 0, not the box lower bound). */
void nxo_particle_index(const double* x, const nxo_box* box, const int grid[3], int idx[3], double frac[3]) {
    for (int d = 0; d < 3; ++d) {
        double recip = 1 / box->L[d];                  /* :19-30 */
        double t = x[d] * recip;
        t = (t - floor(t)) * grid[d];
        int ti = (int)t;
        frac[d] = t - ti;
        idx[d] = ti % grid[d];
    }
}

#define PME_ORDER 5

/* NonbondedForce.cpp:220-260 */
void nxo_pme_spread(int n, const double* pos, const double* q, const nxo_box* box, const int grid[3], double* cg) {
    const int K0 = grid[0], K1 = grid[1], K2 = grid[2];
    memset(cg, 0, sizeof(double) * 2 * (size_t)K0 * K1 * K2);
    for (int i = 0; i < n; ++i) {
        int id[3]; double fr[3], th[3][PME_ORDER], dth[3][PME_ORDER];
        nxo_particle_index(pos + 3 * i, box, grid, id, fr);
        for (int d = 0; d < 3; ++d) nxo_bspline_theta(fr[d], PME_ORDER, th[d], dth[d]);
        for (int ix = 0; ix < PME_ORDER; ++ix) {
            int xi = (id[0] + ix) % K0;
            for (int iy = 0; iy < PME_ORDER; ++iy) {
                int yi = (id[1] + iy) % K1;
                for (int iz = 0; iz < PME_ORDER; ++iz) {
                    int zi = (id[2] + iz) % K2;
                    size_t index = (size_t)xi * K1 * K2 + (size_t)yi * K2 + zi;
                    cg[2 * index] += q[i] * th[0][ix] * th[1][iy] * th[2][iz];
                }
            }
        }
    }
}

/* One complex DFT of length n along a strided line, in place.
 * Mixed-radix decimation in time, recursive on the smallest prime factor;
 * a generic O(p^2) butterfly covers any prime. */
static void dft_rec(int n, const double* in, int istride, double* out, int sign, double* scratch) {
    if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }
    int p = 2;
    while (n % p) { ++p; if (p * p > n) { p = n; break; } }
    int m = n / p;
    /* sub-transforms of the p decimated sequences into out[s*m .. ] */
    for (int s = 0; s < p; ++s)
        dft_rec(m, in + 2 * s * istride, istride * p, out + 2 * s * m, sign, scratch);
    /* combine: X[k + j*m] = sum_s W_n^{s(k+j m)} Y_s[k] */
    for (int k = 0; k < m; ++k) {
        for (int s = 0; s < p; ++s) {
            scratch[2 * s] = out[2 * (s * m + k)];
            scratch[2 * s + 1] = out[2 * (s * m + k) + 1];
        }
        for (int j = 0; j < p; ++j) {
            int kk = k + j * m;
            double re = 0.0, im = 0.0;
            for (int s = 0; s < p; ++s) {
                long e = ((long)s * kk) % n;
                double ang = sign * 2.0 * NXO_PI * (double)e / (double)n;
                double c = cos(ang), sn = sin(ang);
                re += scratch[2 * s] * c - scratch[2 * s + 1] * sn;
                im += scratch[2 * s] * sn + scratch[2 * s + 1] * c;
            }
            out[2 * kk] = re;
            out[2 * kk + 1] = im;
        }
    }
}

void nxo_fft1d(int n, double* data, int stride, int sign) {
    if (n <= 0) return;
    double* in = (double*)malloc(sizeof(double) * 2 * n);
    double* out = (double*)malloc(sizeof(double) * 2 * n);
    double* scratch = (double*)malloc(sizeof(double) * 2 * n);
    for (int i = 0; i < n; ++i) { in[2 * i] = data[2 * (size_t)i * stride]; in[2 * i + 1] = data[2 * (size_t)i * stride + 1]; }
    dft_rec(n, in, 1, out, sign, scratch);
    for (int i = 0; i < n; ++i) { data[2 * (size_t)i * stride] = out[2 * i]; data[2 * (size_t)i * stride + 1] = out[2 * i + 1]; }
    free(in); free(out); free(scratch);
}

/* 3D unnormalised C2C on a row-major [nx][ny][nz] complex array. */
void nxo_fft3d(int nx, int ny, int nz, double* a, int sign) {
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y) nxo_fft1d(nz, a + 2 * ((size_t)x * ny * nz + (size_t)y * nz), 1, sign);
    for (int x = 0; x < nx; ++x)
        for (int z = 0; z < nz; ++z) nxo_fft1d(ny, a + 2 * ((size_t)x * ny * nz + z), nz, sign);
    for (int y = 0; y < ny; ++y)
        for (int z = 0; z < nz; ++z) nxo_fft1d(nx, a + 2 * ((size_t)y * nz + z), ny * nz, sign);
}

/* NonbondedForce.cpp:323-388; returns 0.5 * sum (the value that :387 assigns). */
double nxo_pme_convolve(double* cg, const nxo_box* box, const int grid[3], double alpha) {
    const int K0 = grid[0], K1 = grid[1], K2 = grid[2];
    double* mod[3];
    for (int d = 0; d < 3; ++d) { mod[d] = (double*)malloc(sizeof(double) * grid[d]); nxo_bspline_moduli(grid[d], PME_ORDER, mod[d]); }
    double rv[3] = {1 / box->L[0], 1 / box->L[1], 1 / box->L[2]};
    double one_4pi_eps = NXO_ONE_4PI_EPS0 / 1.0;
    double factor = NXO_PI * NXO_PI / (alpha * alpha);
    double boxfactor = NXO_PI * box->L[0] * box->L[1] * box->L[2];
    double esum = 0.0;
    int mkx = (K0 + 1) / 2, mky = (K1 + 1) / 2, mkz = (K2 + 1) / 2;
    for (int kx = 0; kx < K0; ++kx) {
        double mx = (kx < mkx) ? kx : (kx - K0);
        double mhx = mx * rv[0];
        double bx = boxfactor * mod[0][kx];
        for (int ky = 0; ky < K1; ++ky) {
            double my = (ky < mky) ? ky : (ky - K1);
            double mhy = my * rv[1];
            double by = mod[1][ky];
            for (int kz = 0; kz < K2; ++kz) {
                if (kx == 0 && ky == 0 && kz == 0) continue;
                double mz = (kz < mkz) ? kz : (kz - K2);
                double mhz = mz * rv[2];
                size_t index = (size_t)kx * K1 * K2 + (size_t)ky * K2 + kz;
                double d1 = cg[2 * index], d2 = cg[2 * index + 1];
                double m2 = mhx * mhx + mhy * mhy + mhz * mhz;
                double bz = mod[2][kz];
                double denom = m2 * bx * by * bz;
                double eterm = one_4pi_eps * exp(-factor * m2) / denom;
                cg[2 * index] = d1 * eterm;
                cg[2 * index + 1] = d2 * eterm;
                esum += eterm * (d1 * d1 + d2 * d2);
            }
        }
    }
    for (int d = 0; d < 3; ++d) free(mod[d]);
    return 0.5 * esum;
}

/* NonbondedForce.cpp:407-461 */
void nxo_pme_gather(int n, const double* pos, const double* q, const nxo_box* box, const int grid[3], const double* cg, double* f) {
    const int K0 = grid[0], K1 = grid[1], K2 = grid[2];
    for (int i = 0; i < n; ++i) {
        int id[3]; double fr[3], th[3][PME_ORDER], dth[3][PME_ORDER];
        nxo_particle_index(pos + 3 * i, box, grid, id, fr);
        for (int d = 0; d < 3; ++d) nxo_bspline_theta(fr[d], PME_ORDER, th[d], dth[d]);
        double fx = 0.0, fy = 0.0, fz = 0.0;
        for (int ix = 0; ix < PME_ORDER; ++ix) {
            int xi = (id[0] + ix) % K0;
            double tx = th[0][ix], dtx = dth[0][ix];
            for (int iy = 0; iy < PME_ORDER; ++iy) {
                int yi = (id[1] + iy) % K1;
                double ty = th[1][iy], dty = dth[1][iy];
                for (int iz = 0; iz < PME_ORDER; ++iz) {
                    int zi = (id[2] + iz) % K2;
                    double tz = th[2][iz], dtz = dth[2][iz];
                    size_t index = (size_t)xi * K1 * K2 + (size_t)yi * K2 + zi;
                    double g = cg[2 * index];
                    fx += dtx * ty * tz * g;
                    fy += tx * dty * tz * g;
                    fz += tx * ty * dtz * g;
                }
            }
        }
        f[3 * i + 0] -= q[i] * (fx * K0 * (1 / box->L[0]));
        f[3 * i + 1] -= q[i] * (fy * K1 * (1 / box->L[1]));
        f[3 * i + 2] -= q[i] * (fz * K2 * (1 / box->L[2]));
    }
}

/* NonbondedForce.cpp:467-499 (initializeErfcTable is dead work, omitted). */
double nxo_pme_recip(int n, const double* pos, const double* q, const nxo_box* box, const int grid[3], double alpha, double* f) {
    size_t G = (size_t)grid[0] * grid[1] * grid[2];
    double* cg = (double*)malloc(sizeof(double) * 2 * G);
    nxo_pme_spread(n, pos, q, box, grid, cg);
    nxo_fft3d(grid[0], grid[1], grid[2], cg, -1);
    double e = nxo_pme_convolve(cg, box, grid, alpha);
    nxo_fft3d(grid[0], grid[1], grid[2], cg, +1);
    nxo_pme_gather(n, pos, q, box, grid, cg, f);
    free(cg);
    return e;
}

/* NonbondedForce.cpp:515-527 (q is unchanged by the modifier :504-513). */
double nxo_self_energy(int n, const double* q, double alpha) {
    double t = 0.0;
    for (int i = 0; i < n; ++i) t -= NXO_ONE_4PI_EPS0 * q[i] * q[i] * alpha / NXO_SQRT_PI;
    return t;
}

/* NonbondedForce.cpp:530-585 over a given pair list (i = first, j = second). */
double nxo_direct(int n, const double* pos, const double* q, const double* sig, const double* eps,
                  const nxo_box* box, const int* pairs, int64_t np, double alpha, double cutoff,
                  int use_switch, double rsw, int quirk, double* f, double* e_coul, double* e_lj) {
    (void)n;
    double ev = 0.0, ec = 0.0;
    double swv = 1.0, swd = 0.0;
    for (int64_t p = 0; p < np; ++p) {
        int ii = pairs[2 * p], jj = pairs[2 * p + 1];
        double dR[3];
        nxo_min_image(pos + 3 * ii, pos + 3 * jj, box, dR);
        double r = sqrt(dR[0] * dR[0] + dR[1] * dR[1] + dR[2] * dR[2]);
        double ir = 1.0 / r;
        if (!quirk) { swv = 1.0; swd = 0.0; }
        if (use_switch && r > rsw) {
            double t = (r - rsw) / (cutoff - rsw);
            swv = 1 + t * t * t * (-10 + t * (15 - t * 6));
            swd = t * t * (-30 + t * (60 - t * 30)) / (cutoff - rsw);
        }
        double ar = alpha * r;
        double dEdR = NXO_ONE_4PI_EPS0 * q[ii] * q[jj] * ir * ir * ir;
        dEdR *= (erfc(ar) + 2 * ar * exp(-ar * ar) / NXO_SQRT_PI);
        double s = 0.5 * sig[ii] + 0.5 * sig[jj];
        double s2 = s * ir;
        s2 *= s2;
        double s6 = s2 * s2 * s2;
        double ep = (2.0 * sqrt(eps[ii])) * (2.0 * sqrt(eps[jj]));
        dEdR += swv * ep * (12.0 * s6 - 6.0) * s6 * ir * ir;
        double vdw = ep * (s6 - 1.0) * s6;
        if (use_switch) { dEdR -= vdw * swd * ir; vdw *= swv; }
        for (int k = 0; k < 3; ++k) { double fk = dR[k] * dEdR; f[3 * ii + k] += fk; f[3 * jj + k] -= fk; }
        ec += NXO_ONE_4PI_EPS0 * q[ii] * q[jj] * ir * erfc(ar);
        ev += vdw;
    }
    *e_coul = ec;
    *e_lj = ev;
    return ec + ev;
}

/* NonbondedForce.cpp:590-630; returns the total that is subtracted. */
double nxo_exclusion_correction(int n, const double* pos, const double* q, const nxo_box* box,
                                const int* off, const int* idx, double alpha, double* f) {
    double tot = 0.0;
    if (!off) return 0.0;
    for (int i = 0; i < n; ++i)
        for (int a = off[i]; a < off[i + 1]; ++a) {
            int j = idx[a];
            if (j <= i) continue;
            double dR[3];
            nxo_min_image(pos + 3 * i, pos + 3 * j, box, dR);
            double r = sqrt(dR[0] * dR[0] + dR[1] * dR[1] + dR[2] * dR[2]);
            double ir = 1.0 / r;
            double ar = alpha * r;
            double ee;
            if (erf(ar) > 1e-6) {
                double dEdR = NXO_ONE_4PI_EPS0 * q[i] * q[j] * ir * ir * ir;
                dEdR = dEdR * (erf(ar) - 2 * ar * exp(-ar * ar) / NXO_SQRT_PI);
                for (int k = 0; k < 3; ++k) { double fk = dR[k] * dEdR; f[3 * i + k] -= fk; f[3 * j + k] += fk; }
                ee = NXO_ONE_4PI_EPS0 * q[i] * q[j] * ir * erf(ar);
            } else {
                ee = alpha * NXO_TWO_OVER_SQRT_PI * NXO_ONE_4PI_EPS0 * q[i] * q[j];
            }
            tot += ee;
        }
    return tot;
}

/* Engine::CalculateForces (core/baseline/src/Engine.cpp:75-95) with the
 * nonbonded composition of NonbondedForce.cpp:633-647. */
int nxo_compute_forces(const nxo_system* s, const double* pos, int mask, double* f, nxo_energies* e) {
    memset(e, 0, sizeof(*e));
    memset(f, 0, sizeof(double) * 3 * (size_t)s->n);
    double pe = 0.0;
    if (mask & NXO_BOND) { e->bond = nxo_bonds(s->n_bonds, s->bonds, pos, &s->box, f); pe += e->bond; }
    if (mask & NXO_ANGLE) { e->angle = nxo_angles(s->n_angles, s->angles, pos, &s->box, f); pe += e->angle; }
    if (mask & NXO_TORSION) { e->torsion = nxo_torsions(s->n_torsions, s->torsions, pos, &s->box, f); pe += e->torsion; }
    e->potential = e->bond + e->angle + e->torsion;
    e->reference_pe = pe;
    if (mask & NXO_NONBONDED) {
        int grid[3];
        if (s->grid[0] > 0) { grid[0] = s->grid[0]; grid[1] = s->grid[1]; grid[2] = s->grid[2]; }
        else nxo_grid_rule(&s->box, grid);
        double alpha = nxo_alpha(s->cutoff, s->ewald_tol);
        double* fn = (double*)calloc(3 * (size_t)s->n, sizeof(double));
        e->recip = nxo_pme_recip(s->n, pos, s->q, &s->box, grid, alpha, fn);
        e->self = nxo_self_energy(s->n, s->q, alpha);
        int64_t cap = 1 << 20;
        int* pairs = (int*)malloc(sizeof(int) * 2 * cap);
        int64_t np = nxo_pairs_voxel(s->n, pos, &s->box, s->excl_off, s->excl_idx, s->cutoff, pairs, cap);
        if (np > cap) {
            free(pairs);
            cap = np;
            pairs = (int*)malloc(sizeof(int) * 2 * cap);
            np = nxo_pairs_voxel(s->n, pos, &s->box, s->excl_off, s->excl_idx, s->cutoff, pairs, cap);
        }
        nxo_direct(s->n, pos, s->q, s->sig, s->eps, &s->box, pairs, np, alpha, s->cutoff, s->use_switch,
                   s->switch_distance, s->switch_quirk, fn, &e->coulomb, &e->lj);
        free(pairs);
        e->exclusion = nxo_exclusion_correction(s->n, pos, s->q, &s->box, s->excl_off, s->excl_idx, alpha, fn);
        for (size_t k = 0; k < 3 * (size_t)s->n; ++k) f[k] += fn[k];
        free(fn);
        /* the reciprocal energy is assigned over the accumulated PE (NonbondedForce.cpp:387) */
        e->reference_pe = e->recip + e->self + (e->coulomb + e->lj) - e->exclusion;
        e->potential += e->recip + e->self + e->coulomb + e->lj - e->exclusion;
    }
    return 0;
}

/* core/baseline/src/VerletIntegration.cpp:37-45 and :49-89 (dt0 = dt1 = dt). */
void nxo_leapfrog(int n, double* pos, double* vel, const double* f, const double* mass, double dt, const nxo_box* box) {
    double dtPos = dt, dtVel = 0.5 * (dt + dt);
    for (int i = 0; i < n; ++i) {
        double im = (mass[i] == 0.0) ? 0.0 : 1.0 / mass[i];
        if (im == 0.0) continue;
        for (int j = 0; j < 3; ++j) {
            vel[3 * i + j] += im * f[3 * i + j] * dtVel;
            double x = pos[3 * i + j] + vel[3 * i + j] * dtPos;
            if (x < box->lb[j]) x += box->L[j];
            else if (x > box->ub[j]) x -= box->L[j];
            pos[3 * i + j] = x;
        }
    }
}

/* core/baseline/src/KineticEnergy.cpp:21-41 */
double nxo_kinetic(int n, const double* vel, const double* mass, const double* f, double dt) {
    const double ts = 0.5 * dt;
    double tot = 0.0;
    for (int i = 0; i < n; ++i) {
        double v[3] = {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
        if (mass[i] > 0) {
            double im = 1.0 / mass[i];
            double a0 = f[3 * i] * im * ts, a1 = f[3 * i + 1] * im * ts, a2 = f[3 * i + 2] * im * ts;
            v[0] += a0; v[1] += a1; v[2] += a2;
        }
        tot += 0.5 * mass[i] * dot3(v, v);
    }
    return tot;
}

/* Engine::RunSimulation (Engine.cpp:162-188) without file output. */
int nxo_simulate(const nxo_system* s, double* pos, double* vel, int mask, int nsteps, double dt, int interval,
                 double* eo, double* f_last) {
    double* f = (double*)malloc(sizeof(double) * 3 * (size_t)s->n);
    int k = 0;
    for (int step = 0; step < nsteps; ++step) {
        nxo_energies e;
        nxo_compute_forces(s, pos, mask, f, &e);
        if (((step + 1) % interval) == 0 || step == 0) {
            double ke = nxo_kinetic(s->n, vel, s->mass, f, dt);
            if (eo) { eo[3 * k] = ke; eo[3 * k + 1] = e.reference_pe; eo[3 * k + 2] = e.reference_pe + ke; }
            ++k;
        }
        nxo_leapfrog(s->n, pos, vel, f, s->mass, dt, &s->box);
    }
    if (f_last) memcpy(f_last, f, sizeof(double) * 3 * (size_t)s->n);
    free(f);
    return k;
}

// nl.hip -- neighbour-list build on the device.
//
// Replaces obtainNeighborList / VoxelHash (core/baseline/src/Forces/
// NeighborList.cpp:12-154), which rebuilds a std::map voxel hash every step.
// Here, per rebuild:
//   1. nl_bin      : counting sort of atoms into xy columns (one atomic per
//                    distinct column in a wavefront);
//   2. nl_scan     : per-system exclusive scan -> slot ranges (columns padded
//                    to 8-atom clusters) and super-cluster ranges; clears the
//                    bin counters for the next build;
//   3. nl_scatter  : atoms into their column's slot range;
//   4. nl_columns  : per column, sort by (z, atom id) in LDS -> deterministic
//                    slot order; cluster bounding boxes; build positions in
//                    slot order; the slot-order positions, charges and LJ
//                    parameters the force kernels read, and each atom's build
//                    position for the displacement test;
//   5. nl_pairlist : one block per super-cluster (8 clusters of 8 atoms):
//                    j-clusters of nearby columns whose bounding box is within
//                    rlist of an i-cluster (half list: j >= i), then per tile
//                    the 64 atom pairs, keeping only tiles with a pair within
//                    rlist; an 8-bit mask of interacting i-clusters per entry,
//                    and 64-bit exclusion masks for tiles holding excluded pairs.
// The pair kernel applies r^2 < rc^2 exactly per pair, so the evaluated pair
// set is the reference's set whatever the buffer (skin) is.
//
// Every kernel first checks the system's rebuild flag, which the integrator
// raises when an atom has moved more than skin/2 since the last build; the
// last kernel (nl_waves) cuts the chunk ranges of the pair kernel's
// wavefronts, and the kernel launched before the pair kernel clears the flag
// (nl_clear_flag when no pair kernel follows).  Between builds the integrator
// keeps the slot-order positions current.
#include "nx_device.h"
#include "nx_engine.h"

namespace nx {

constexpr int kColSortCap = 2048;     // atoms per column sorted in LDS
constexpr int kListCap = kSuperListCap;   // pair-list entries per super-cluster (LDS)
constexpr int kExclCap = 2048;        // exclusion triples per super-cluster (LDS)

// Column of an atom from its wrapped position.
__device__ __forceinline__ void wrapped(const SysDev& d, double x, double y, double z, double* w) {
    double p[3] = {x, y, z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double v = p[k] - floor(p[k] * d.invL[k]) * d.L[k];
        if (!(v >= 0.0)) v = 0.0;              // NaN or rounding below 0
        if (v >= d.L[k]) v = 0.0;              // rounding up to L wraps to 0
        w[k] = v;
    }
}

// Adds one to counter[key] for every active lane of the wavefront with one
// atomic per distinct key (neighbouring atoms mostly share a column, so a
// wavefront touches few columns); returns the counter's old value plus the
// lane's rank among the lanes that share its key.
__device__ __forceinline__ int wave_keyed_increment(int* counter, int key) {
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    bool done = false;
    int res = 0;
    for (;;) {
        const unsigned long long pending = __ballot(!done);
        if (pending == 0ull) break;
        const int leader = __builtin_ctzll(pending);
        const int k = __shfl(key, leader);
        const bool mine = !done && key == k;
        const unsigned long long grp = __ballot(mine);
        int base = 0;
        if (lane == leader) base = atomicAdd(&counter[k], __popcll(grp));
        base = __shfl(base, leader);
        if (mine) {
            res = base + __popcll(grp & below);
            done = true;
        }
    }
    return res;
}

// Column of every atom and the column counts (into col_tmp, zero between builds).
__global__ void nl_bin(const SysDev* sd, SysCount* sc, const int* atom_sys, const double* X, const double* Y,
                       const double* Z, int* atom_col, int* col_tmp, int n_atoms_total) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n_atoms_total) return;
    int s = atom_sys[a];
    if (!sc[s].rebuild) return;
    const SysDev& d = sd[s];
    double w[3];
    wrapped(d, X[a], Y[a], Z[a], w);
    int cx = (int)(w[0] / d.colw[0]);
    int cy = (int)(w[1] / d.colw[1]);
    cx = cx < 0 ? 0 : (cx >= d.ncx ? d.ncx - 1 : cx);
    cy = cy < 0 ? 0 : (cy >= d.ncy ? d.ncy - 1 : cy);
    int col = d.col_off + cx * d.ncy + cy;
    atom_col[a] = col;
    wave_keyed_increment(col_tmp, col);
}

// One block per system: scan padded column counts; col_tmp -> col_count and
// back to zero.
__global__ void nl_scan(const SysDev* sd, SysCount* sc, int* col_tmp, int* col_count, int* col_slot,
                        int* col_super) {
    int s = blockIdx.x;
    if (!sc[s].rebuild) return;
    const SysDev& d = sd[s];
    const int nc = d.ncx * d.ncy;
    // the build now happens: a stale-list mark of an earlier step without the
    // list kernels is cleared before the kernels after this one read the flags
    if (threadIdx.x == 0 && (sc[s].overflow & kStaleList)) sc[s].overflow &= ~kStaleList;
    __shared__ int tot_slot, tot_sup, carry_slot, carry_sup;
    __shared__ int ws[1024], wsup[1024];
    if (threadIdx.x == 0) { carry_slot = 0; carry_sup = 0; }
    __syncthreads();
    for (int base = 0; base < nc; base += blockDim.x) {
        int c = base + threadIdx.x;
        int cnt = 0;
        if (c < nc) {
            cnt = col_tmp[d.col_off + c];
            col_count[d.col_off + c] = cnt;
            col_tmp[d.col_off + c] = 0;
        }
        int ps = (cnt + 7) & ~7;
        int pu = (cnt + 63) >> 6;
        ws[threadIdx.x] = ps;
        wsup[threadIdx.x] = pu;
        __syncthreads();
        // Hillis-Steele inclusive scan
        for (int o = 1; o < (int)blockDim.x; o <<= 1) {
            int v1 = threadIdx.x >= (unsigned)o ? ws[threadIdx.x - o] : 0;
            int v2 = threadIdx.x >= (unsigned)o ? wsup[threadIdx.x - o] : 0;
            __syncthreads();
            ws[threadIdx.x] += v1;
            wsup[threadIdx.x] += v2;
            __syncthreads();
        }
        if (c < nc) {
            col_slot[d.col_off + c] = d.slot_off + carry_slot + ws[threadIdx.x] - ps;
            col_super[d.col_off + c] = d.super_off + carry_sup + wsup[threadIdx.x] - pu;
        }
        __syncthreads();
        if (threadIdx.x == blockDim.x - 1) { carry_slot += ws[threadIdx.x]; carry_sup += wsup[threadIdx.x]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        tot_slot = carry_slot;
        tot_sup = carry_sup;
        int ovf = (tot_slot > d.slot_cap) || (tot_sup > d.super_cap);
        sc[s].n_slots = ovf ? 0 : tot_slot;
        sc[s].n_clusters = ovf ? 0 : tot_slot / 8;
        sc[s].n_super = ovf ? 0 : tot_sup;
        sc[s].list_count = 0;
        sc[s].emask_count = 0;
        sc[s].n_chunks = 0;
        if (ovf) sc[s].overflow |= 1;
    }
}

__global__ void nl_scatter(const SysDev* sd, SysCount* sc, const int* atom_sys, const int* atom_col,
                           const int* col_slot, int* col_fill, int* slot_atom_tmp, int n_atoms_total) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n_atoms_total) return;
    int s = atom_sys[a];
    if (!sc[s].rebuild || (sc[s].overflow & 1)) return;
    int col = atom_col[a];
    int k = wave_keyed_increment(col_fill, col);
    int slot = col_slot[col] + k;
    if (slot < sd[s].slot_off + sd[s].slot_cap) slot_atom_tmp[slot] = a;
}

// One block (256 threads) per column: deterministic z order, cluster boxes,
// super items, and the slot-order data of the force kernels (pos: x, y, z, q;
// kPadPos coordinates and q = 0 on padding slots; lj: sigma/2, 2 sqrt(eps)).
template <class P>
__global__ __launch_bounds__(256) void nl_columns(const SysDev* sd, SysCount* sc, const int* col_sys,
                                                  const int* col_count, int* col_fill, const int* col_slot,
                                                  const int* col_super, const int* slot_atom_tmp, const double* X,
                                                  const double* Y, const double* Z, const double* Q,
                                                  const double* sig, const double* eps, int* atom_of_slot,
                                                  int* slot_of_atom, float4* cl_ctr, float4* cl_half,
                                                  int4* super_item, float* slot_zkey, float4* slot_wpos,
                                                  typename P::real4* pos, typename P::real2* lj, double* XB,
                                                  double* YB, double* ZB) {
    using real = typename P::real;
    const int col = blockIdx.x;
    const int s = col_sys[col];
    if (!sc[s].rebuild || (sc[s].overflow & 1)) return;
    if (threadIdx.x == 0) col_fill[col] = 0;          // the scatter is done: zero for the next build
    const SysDev& d = sd[s];
    const int cnt = col_count[col];
    const int first = col_slot[col];
    __shared__ double key[kColSortCap];
    __shared__ int ids[kColSortCap];
    if (cnt > kColSortCap) {
        if (threadIdx.x == 0) atomicOr(&sc[s].overflow, 2);
        return;
    }
    int np2 = 1;
    while (np2 < cnt) np2 <<= 1;
    for (int i = threadIdx.x; i < np2; i += blockDim.x) {
        if (i < cnt) {
            int a = slot_atom_tmp[first + i];
            double w[3];
            wrapped(d, X[a], Y[a], Z[a], w);
            key[i] = w[2];
            ids[i] = a;
        } else {
            key[i] = 1e300;
            ids[i] = 0x7fffffff;
        }
    }
    __syncthreads();
    // bitonic sort by (z, atom id)
    for (int k = 2; k <= np2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < np2; i += blockDim.x) {
                int l = i ^ j;
                if (l > i) {
                    bool up = (i & k) == 0;
                    bool gt = (key[i] > key[l]) || (key[i] == key[l] && ids[i] > ids[l]);
                    if (gt == up) {
                        double tk = key[i]; key[i] = key[l]; key[l] = tk;
                        int ti = ids[i]; ids[i] = ids[l]; ids[l] = ti;
                    }
                }
            }
            __syncthreads();
        }
    }
    const int padded = (cnt + 7) & ~7;
    for (int i = threadIdx.x; i < padded; i += blockDim.x) {
        int slot = first + i;
        if (i < cnt) {
            const int a = ids[i];
            atom_of_slot[slot] = a;
            slot_of_atom[a] = slot;
            // round down so that a key never exceeds the true z (slab binning stays monotone)
            float zk = (float)key[i];
            if ((double)zk > key[i]) zk = nextafterf(zk, -1.0f);
            slot_zkey[slot] = zk;
            double w[3];
            wrapped(d, X[a], Y[a], Z[a], w);
            slot_wpos[slot] = make_float4((float)w[0], (float)w[1], (float)w[2], 1.0f);
            // the build position is the wrapped one; the mixed path's slot
            // positions continue from it until the next build, the FP64 path's
            // are the positions themselves (the reference's arithmetic)
            typename P::real4 p;
            p.x = (real)(P::mixed ? w[0] : X[a]);
            p.y = (real)(P::mixed ? w[1] : Y[a]);
            p.z = (real)(P::mixed ? w[2] : Z[a]);
            p.w = (real)Q[a];
            pos[slot] = p;
            typename P::real2 l;
            l.x = (real)(0.5 * sig[a]);
            l.y = (real)(2.0 * sqrt(eps[a]));
            lj[slot] = l;
            XB[a] = w[0];
            YB[a] = w[1];
            ZB[a] = w[2];
        } else {
            atom_of_slot[slot] = -1;
            const float nan = __int_as_float(0x7fc00000);     // compares false in every distance test
            slot_wpos[slot] = make_float4(nan, nan, nan, 0.f);
            typename P::real4 p;
            p.x = p.y = p.z = (real)kPadPos;
            p.w = 0;
            pos[slot] = p;
            typename P::real2 l;
            l.x = 0;
            l.y = 0;
            lj[slot] = l;
        }
    }
    // cluster bounding boxes in wrapped coordinates
    const int ncl = padded >> 3;
    for (int c = threadIdx.x; c < ncl; c += blockDim.x) {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int t = 0; t < 8; ++t) {
            int i = c * 8 + t;
            if (i >= cnt) break;
            int a = ids[i];
            double w[3];
            wrapped(d, X[a], Y[a], Z[a], w);
            for (int k = 0; k < 3; ++k) { lo[k] = fmin(lo[k], w[k]); hi[k] = fmax(hi[k], w[k]); }
        }
        int gcl = (first >> 3) + c;
        float4 ctr, half;
        ctr.x = (float)(0.5 * (lo[0] + hi[0]));
        ctr.y = (float)(0.5 * (lo[1] + hi[1]));
        ctr.z = (float)(0.5 * (lo[2] + hi[2]));
        ctr.w = 0.f;
        half.x = (float)(0.5 * (hi[0] - lo[0])) + 1e-5f;
        half.y = (float)(0.5 * (hi[1] - lo[1])) + 1e-5f;
        half.z = (float)(0.5 * (hi[2] - lo[2])) + 1e-5f;
        half.w = 0.f;
        cl_ctr[gcl] = ctr;
        cl_half[gcl] = half;
    }
    // super-cluster items of this column
    const int nsup = (cnt + 63) >> 6;
    const int sup0 = col_super[col];
    for (int u = threadIdx.x; u < nsup; u += blockDim.x) {
        int4 it;
        it.x = (first >> 3) + u * 8;                 // first global cluster
        it.y = min(8, ncl - u * 8);                  // clusters in the item
        it.z = 0;
        it.w = 0;
        super_item[sup0 + u] = it;
    }
}

__device__ __forceinline__ float bbox_gap2(float4 ca, float4 ha, float4 cb, float4 hb, float Lx, float Ly, float Lz) {
    float dx = fabsf(ca.x - cb.x), dy = fabsf(ca.y - cb.y), dz = fabsf(ca.z - cb.z);
    dx = fminf(dx, Lx - dx);
    dy = fminf(dy, Ly - dy);
    dz = fminf(dz, Lz - dz);
    dx = fmaxf(0.f, dx - ha.x - hb.x);
    dy = fmaxf(0.f, dy - ha.y - hb.y);
    dz = fmaxf(0.f, dz - ha.z - hb.z);
    return dx * dx + dy * dy + dz * dz;
}

// The periodic image (nx, ny, nz in -1..1) of cluster b as seen from cluster a,
// from their wrapped build-time boxes, as the code (nx+1) + 3 (ny+1) + 9 (nz+1):
// every atom pair (p in a, q in b) then has p - q - n L as its nearest image
// vector until the next build, because |c_a - c_b - n L| + h_a + h_b + slack
// < L/2 on every axis (slack: the atoms' motion until the build, and
// round-off).  -1 when the boxes do not guarantee it.
constexpr int kImgMixed = 27;      // the entry's tiles need different images
__device__ __forceinline__ int tile_image(float4 ca, float4 ha, float4 cb, float4 hb, float Lx, float Ly, float Lz,
                                          float iLx, float iLy, float iLz, float slack) {
    const float dx = ca.x - cb.x, dy = ca.y - cb.y, dz = ca.z - cb.z;
    const float nx = rintf(dx * iLx), ny = rintf(dy * iLy), nz = rintf(dz * iLz);
    const bool ok = fabsf(dx - nx * Lx) + ha.x + hb.x + slack < 0.5f * Lx &&
                    fabsf(dy - ny * Ly) + ha.y + hb.y + slack < 0.5f * Ly &&
                    fabsf(dz - nz * Lz) + ha.z + hb.z + slack < 0.5f * Lz;
    return ok ? (int)nx + 3 * (int)ny + 9 * (int)nz + 13 : -1;
}
// The list word's image bits (two per axis: 0 none, 1 +L, 2 -L) of an image code.
__device__ __forceinline__ unsigned image_bits(int code) {
    const int n[3] = {code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1};
    unsigned b = 0;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) b |= (n[ax] == 1 ? 1u : (n[ax] == -1 ? 2u : 0u)) << (2 * ax);
    return b;
}

// Position in increasing index order of the k-th column of a periodic window of
// nwin columns centred on c0 (nwin == nc: every column).  Enumerating windows
// this way makes the cluster indices of the candidates increase, so the list
// comes out sorted without a sort.
__device__ __forceinline__ int window_col(int k, int c0, int D, int nc, int nwin) {
    if (nwin == nc) return k;
    const int st = ((c0 - D) % nc + nc) % nc;
    const int wrapped_front = st + nwin - nc;          // columns past nc - 1, wrapped to 0..
    if (wrapped_front > 0) return k < wrapped_front ? k : st + (k - wrapped_front);
    return st + k;
}

constexpr int kPlThreads = 256;       // threads per super-cluster in nl_pairlist

// Exclusive prefix sum over the kPlThreads threads of a block (thread order);
// `total` receives the block sum.  Every thread must call it.
__device__ __forceinline__ int block_excl_scan(int v, int* wsum, int& total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kPlThreads / 64; ++w) {
        const int x = wsum[w];
        off += (w < wv) ? x : 0;
        tot += x;
    }
    __syncthreads();
    total = tot;
    return off + incl - v;
}

// One block of kPlThreads per super-cluster item:
//   1. the columns of the xy window whose rectangle comes within rlist of the
//      super-cluster's box, in increasing column order;
//   2. their clusters as one flat candidate range, one thread per candidate:
//      cluster box against each i-cluster box, ordered compaction; the
//      entry's periodic image (tile_image), when one image holds for all of
//      its tiles, else the per-pair marker;
//   3. per entry and i-cluster, the 64 atom pairs at build positions, one
//      wavefront per entry: a tile is kept only if one of its pairs is within
//      rlist (both atoms move at most skin/2 before the next build, so every
//      pair that can come within rc is in a kept tile); ordered compaction;
//   4. exclusion masks for the kept tiles.
__global__ __launch_bounds__(kPlThreads) void nl_pairlist(const SysDev* sd, SysCount* sc, const int* super_sys,
                                                          int4* super_item, const int* col_count,
                                                          const int* col_slot, const float4* cl_ctr,
                                                          const float4* cl_half, const float4* slot_wpos,
                                                          const int* atom_of_slot, const int* slot_of_atom,
                                                          const int* excl_off, const int* excl_idx, uint2* list,
                                                          unsigned long long* emask, int4* chunks, int* chunk_cost,
                                                          int n_items_total) {
    const int item = xcd_remap(blockIdx.x, gridDim.x);
    if (item >= n_items_total) return;
    const int s = super_sys[item];
    const SysDev& d = sd[s];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    __shared__ int go;
    // one read of the flags per block: another block may raise overflow meanwhile
    if (tid == 0) go = sc[s].rebuild && !sc[s].overflow && item - d.super_off < sc[s].n_super;
    __syncthreads();
    if (!go) return;
    int4 it = super_item[item];
    const int fc = it.x, ncl = it.y;

    __shared__ unsigned int l_ent[kListCap];      // cj | imask << 24
    __shared__ int l_eid[kListCap];
    __shared__ unsigned char l_img[kListCap];     // periodic image code of the entry's j cluster
    __shared__ unsigned int x_rec[kExclCap];      // (entry * 8 + i-cluster) << 6 | pair bit
    __shared__ int n_x, id_base, lbase, cbase;
    __shared__ float4 ic[8], ih[8];
    __shared__ float4 ip[64];                     // i-atom build positions (NaN for padding)
    __shared__ int w_first[kPlThreads], w_pre[kPlThreads];
    __shared__ int wsum[kPlThreads / 64];
    if (tid == 0) n_x = 0;
    if (tid < 8) {
        if (tid < ncl) { ic[tid] = cl_ctr[fc + tid]; ih[tid] = cl_half[fc + tid]; }
        else { ic[tid] = make_float4(1e30f, 1e30f, 1e30f, 0.f); ih[tid] = make_float4(0.f, 0.f, 0.f, 0.f); }
    }
    if (tid < 64) {
        const float nan = __int_as_float(0x7fc00000);
        ip[tid] = ((tid >> 3) < ncl) ? slot_wpos[fc * 8 + tid] : make_float4(nan, nan, nan, 0.f);
    }
    __syncthreads();
    const float Lx = (float)d.L[0], Ly = (float)d.L[1], Lz = (float)d.L[2];
    const float iLx = (float)d.invL[0], iLy = (float)d.invL[1], iLz = (float)d.invL[2];
    const float rl2 = (float)(d.rlist * d.rlist) * 1.0001f;
    // what the tile image test adds to the cluster extents: the motion of two
    // atoms until the next build (skin/2 each) and 1e-3 nm for float round-off
    const float img_slack = (float)(d.rlist - d.rc) + 1e-3f;

    // xy box of the super-cluster
    float xlo = 1e30f, xhi = -1e30f, ylo = 1e30f, yhi = -1e30f;
    for (int k = 0; k < ncl; ++k) {
        xlo = fminf(xlo, ic[k].x - ih[k].x); xhi = fmaxf(xhi, ic[k].x + ih[k].x);
        ylo = fminf(ylo, ic[k].y - ih[k].y); yhi = fmaxf(yhi, ic[k].y + ih[k].y);
    }
    const float scx = 0.5f * (xlo + xhi), scy = 0.5f * (ylo + yhi);
    const float shx = 0.5f * (xhi - xlo), shy = 0.5f * (yhi - ylo);
    const float cwx = (float)d.colw[0], cwy = (float)d.colw[1];

    // window of columns around the super-cluster's own column
    int cx0 = (int)(ic[0].x / cwx), cy0 = (int)(ic[0].y / cwy);
    cx0 = min(max(cx0, 0), d.ncx - 1);
    cy0 = min(max(cy0, 0), d.ncy - 1);
    const int Dx = (int)ceil((d.rlist + 1e-4) / d.colw[0]) + 1;
    const int Dy = (int)ceil((d.rlist + 1e-4) / d.colw[1]) + 1;
    const int nx_c = (2 * Dx + 1 >= d.ncx) ? d.ncx : 2 * Dx + 1;
    const int ny_c = (2 * Dy + 1 >= d.ncy) ? d.ncy : 2 * Dy + 1;
    const int nwin = nx_c * ny_c;

    // 1-2: candidates by cluster boxes
    int ne = 0;                                   // entries so far (the same in every thread)
    for (int w0 = 0; w0 < nwin; w0 += kPlThreads) {
        const int w = w0 + tid;
        int ccl = 0, cfirst = 0;
        if (w < nwin) {
            const int cx = window_col(w / ny_c, cx0, Dx, d.ncx, nx_c);
            const int cy = window_col(w % ny_c, cy0, Dy, d.ncy, ny_c);
            // gap between the column's rectangle (with 1e-3 nm of slack for binning
            // round-off) and the super-cluster's xy box, nearest image
            float gx = fabsf(((float)cx + 0.5f) * cwx - scx), gy = fabsf(((float)cy + 0.5f) * cwy - scy);
            gx = fmaxf(0.f, fminf(gx, Lx - gx) - 0.5f * cwx - 1e-3f - shx);
            gy = fmaxf(0.f, fminf(gy, Ly - gy) - 0.5f * cwy - 1e-3f - shy);
            if (gx * gx + gy * gy < rl2) {
                const int col = d.col_off + cx * d.ncy + cy;
                ccl = (col_count[col] + 7) >> 3;
                cfirst = col_slot[col] >> 3;
            }
        }
        int total;
        const int pre = block_excl_scan(ccl, wsum, total);
        w_pre[tid] = pre;
        w_first[tid] = cfirst;
        __syncthreads();
        for (int b = 0; b < total; b += kPlThreads) {
            const int t = b + tid;
            unsigned int im = 0, img = kImgPerPair;
            int cj = 0;
            if (t < total) {
                int lo = 0, hi = kPlThreads - 1;      // last column whose range starts at or before t
                while (lo < hi) {
                    const int m = (lo + hi + 1) >> 1;
                    if (w_pre[m] <= t) lo = m; else hi = m - 1;
                }
                cj = w_first[lo] + (t - w_pre[lo]);
                const float4 jc = cl_ctr[cj], jh = cl_half[cj];
                int code = -1;
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (k < ncl && cj >= fc + k && bbox_gap2(ic[k], ih[k], jc, jh, Lx, Ly, Lz) < rl2) {
                        im |= 1u << k;
                        const int ck = tile_image(ic[k], ih[k], jc, jh, Lx, Ly, Lz, iLx, iLy, iLz, img_slack);
                        code = (ck < 0 || (code >= 0 && ck != code) || code == kImgMixed) ? kImgMixed : ck;
                    }
                img = (d.pair_image && code >= 0 && code != kImgMixed) ? image_bits(code) : kImgPerPair;
            }
            int cnt;
            const int pos = ne + block_excl_scan(im != 0, wsum, cnt);
            if (im && pos < kListCap) {
                l_ent[pos] = (unsigned)cj | (im << 24);
                l_img[pos] = (unsigned char)img;
            }
            ne += cnt;
        }
        __syncthreads();
    }
    bool overflow = false;
    if (ne > kListCap) { overflow = true; ne = kListCap; }

    // 3: atom pairs, one wavefront per entry, batches of eight entries per
    // wavefront so that their j loads overlap; the kept mask is written in place
    {
        const int ia = lane >> 3, ja = lane & 7;
        float4 pi[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) pi[k] = ip[k * 8 + ia];
        for (int e0 = wv * 8; e0 < ne; e0 += kPlThreads / 8) {
            unsigned int ent[8];
            float4 pj[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                ent[u] = (e0 + u < ne) ? l_ent[e0 + u] : 0u;
                pj[u] = slot_wpos[(ent[u] & 0xffffffu) * 8 + ja];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int cj = (int)(ent[u] & 0xffffffu);
                const unsigned int im = ent[u] >> 24;
                unsigned int keep = 0;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    if (im & (1u << k)) {
                        float dx = pi[k].x - pj[u].x, dy = pi[k].y - pj[u].y, dz = pi[k].z - pj[u].z;
                        dx -= Lx * rintf(dx * iLx);
                        dy -= Ly * rintf(dy * iLy);
                        dz -= Lz * rintf(dz * iLz);
                        // NaN (padding) compares false
                        bool ok = dx * dx + dy * dy + dz * dz < rl2;
                        if (cj == fc + k) ok = ok && ja > ia;
                        if (__ballot(ok)) keep |= 1u << k;
                    }
                }
                if (lane == 0 && e0 + u < ne) l_ent[e0 + u] = (unsigned)cj | (keep << 24);
            }
        }
    }
    __syncthreads();
    {
        int nk = 0;
        for (int b = 0; b < ne; b += kPlThreads) {
            const int e = b + tid;
            const unsigned int v = (e < ne) ? l_ent[e] : 0u;
            const unsigned char vi = (e < ne) ? l_img[e] : (unsigned char)0;
            const bool kept = (v >> 24) != 0u;
            int cnt;
            const int pos = nk + block_excl_scan(kept, wsum, cnt);   // every read precedes every write
            if (kept) { l_ent[pos] = v; l_img[pos] = vi; }
            nk += cnt;
            __syncthreads();
        }
        ne = nk;
    }
    for (int e = tid; e < ne; e += kPlThreads) l_eid[e] = -1;
    __syncthreads();

    // 4: exclusions of the 64 i-atoms that fall in kept tiles, four threads per atom
    {
        const int il = tid & 63, part = tid >> 6;
        const int k = il >> 3, iat = il & 7;
        if (k < ncl) {
            const int si = (fc + k) * 8 + iat;
            const int ai = atom_of_slot[si];
            if (ai >= 0) {
                for (int e = excl_off[ai] + part; e < excl_off[ai + 1]; e += kPlThreads / 64) {
                    const int p = excl_idx[e];
                    const int sp = slot_of_atom[p];
                    const int cj = sp >> 3;
                    if (cj < fc + k) continue;
                    if (cj == fc + k && sp <= si) continue;
                    int lo = 0, hi = ne;              // entries are sorted by cluster
                    while (lo < hi) {
                        const int m = (lo + hi) >> 1;
                        if ((int)(l_ent[m] & 0xffffffu) < cj) lo = m + 1; else hi = m;
                    }
                    if (lo < ne && (int)(l_ent[lo] & 0xffffffu) == cj && ((l_ent[lo] >> 24) >> k & 1u)) {
                        const int t = atomicAdd(&n_x, 1);
                        if (t < kExclCap) x_rec[t] = ((unsigned)(lo * 8 + k) << 6) | (unsigned)(iat * 8 + (sp & 7));
                        l_eid[lo] = 0;      // mark
                    }
                }
            }
        }
    }
    __syncthreads();
    int nx_ = n_x;
    if (nx_ > kExclCap) { overflow = true; nx_ = kExclCap; }
    // mask ids for the entries with exclusions, in entry order
    {
        int nm = 0;
        for (int b = 0; b < ne; b += kPlThreads) {
            const int e = b + tid;
            const bool has = e < ne && l_eid[e] == 0;
            int cnt;
            const int r = block_excl_scan(has, wsum, cnt);
            if (has) l_eid[e] = nm + r;
            nm += cnt;
        }
        if (tid == 0) {
            const int b = atomicAdd(&sc[s].emask_count, nm);
            id_base = (b + nm > d.emask_cap) ? -1 : b;
        }
    }
    __syncthreads();
    if (id_base < 0) overflow = true;
    if (!overflow) {
        for (int e = tid; e < ne; e += kPlThreads) {
            if (l_eid[e] >= 0) {
                const int g = d.emask_off + id_base + l_eid[e];
                for (int k = 0; k < 8; ++k) emask[(size_t)g * 8 + k] = 0ull;
            }
        }
        __syncthreads();
        for (int t = tid; t < nx_; t += kPlThreads) {
            const int ek = (int)(x_rec[t] >> 6), e = ek >> 3, k = ek & 7;
            const int g = d.emask_off + id_base + l_eid[e];
            atomicOr(&emask[(size_t)g * 8 + k], 1ull << (x_rec[t] & 63u));
        }
    }
    const int nch = (ne + kChunkEntries - 1) / kChunkEntries;
    if (tid == 0) {
        const int b = atomicAdd(&sc[s].list_count, ne);
        lbase = (b + ne > d.list_cap) ? -1 : b;
        const int c = atomicAdd(&sc[s].n_chunks, nch);
        cbase = (c + nch > d.chunk_cap) ? -1 : c;
    }
    __syncthreads();
    if (lbase < 0 || cbase < 0) overflow = true;
    if (overflow) {
        if (tid == 0) atomicOr(&sc[s].overflow, 4);
        return;
    }
    for (int e = tid; e < ne; e += kPlThreads) {
        uint2 v;
        v.x = l_ent[e];
        v.y = ((l_eid[e] >= 0) ? (unsigned)(d.emask_off + id_base + l_eid[e]) : kNoMask) |
              ((unsigned)l_img[e] << kImgShift);
        list[d.list_off + lbase + e] = v;
    }
    if (tid == 0) {
        it.z = d.list_off + lbase;
        it.w = ne;
        super_item[item] = it;
    }
    for (int c = tid; c < nch; c += kPlThreads) {
        const int e0 = d.list_off + lbase + c * kChunkEntries;
        const int e1 = d.list_off + lbase + min(ne, (c + 1) * kChunkEntries);
        chunks[d.chunk_off + cbase + c] = make_int4(fc, ncl, e0, e1);
        int cost = kCostChunk;
        for (int e = c * kChunkEntries; e < min(ne, (c + 1) * kChunkEntries); ++e)
            cost += kCostEntry + kCostTile * __popc(l_ent[e] >> 24);
        chunk_cost[d.chunk_off + cbase + c] = cost;
    }
}

// Last kernel of the list work, one block.  After a build in any system of the
// batch: the batch's chunk order (system by system, prefix[]) and, for the NW
// wavefronts of the pair kernel, contiguous chunk ranges of near-equal cost:
// chunk g goes to wavefront floor(NW * (midpoint of g's cost interval) / total
// cost), so wave_start[w] is the first chunk whose midpoint reaches w / NW of
// the total.
constexpr int kWavesLdsSys = 1024;     // systems whose chunk prefix nl_waves keeps in LDS

// With `pool` non-null it also does what the pair kernel's set-up launch does
// (after every read of the flags): clears the rebuild flags, counts the
// builds and resets the pair kernel's claim counters.  In a step without the
// list kernels (`lite`) a raised flag is kept and the system is marked as
// having a stale list instead: the update then freezes it for the step.
// With `info` non-null (a one-system batch) thread 0 also reports to host
// memory what the host needs to choose the next step graphs: the largest
// displacement after the last update (`dmax2`, from the update's slots),
// whether this step builds a list (or, without the list kernels, froze), the
// steps done, and the report count, in one 16-byte store.  The report is a
// forecast aid only: the device freezes a step whose list is due whatever
// the host chose.
__device__ __forceinline__ void waves_pair_begin(SysCount* sc, int nsys, int* pool, int lite, unsigned* step_seq,
                                                 StepInfo* info, float dmax2) {
    const int tid = threadIdx.x;
    if (tid < 8) pool[tid * kPoolStride] = 0;
    for (int s = tid; s < nsys; s += blockDim.x) {
        const int due = sc[s].rebuild;
        if (due) {
            if (lite) {
                sc[s].overflow |= kStaleList;
            } else {
                sc[s].rebuild = 0;
                sc[s].rebuilds += 1;
            }
        }
        if (info && s == 0) {
            const unsigned n = *step_seq + 1u;     // one thread of one block per launch
            *step_seq = n;
            const int flags = (due ? 1 : 0) | (sc[0].overflow << 1);
            *reinterpret_cast<uint4*>(info) =
                make_uint4(n, __float_as_uint(sqrtf(dmax2)), (unsigned)sc[0].steps_done, (unsigned)flags);
        }
    }
}

__global__ __launch_bounds__(1024) void nl_waves(const SysDev* sd, SysCount* sc, int nsys, const int* chunk_cost,
                                                 int* prefix, int* wave_start, int NW, int* pool, int lite,
                                                 unsigned* step_seq, StepInfo* info, unsigned* step_max) {
    __shared__ int any, total;
    __shared__ long long wsum[16];
    __shared__ float dmax_s[1];
    __shared__ int pre_s[kWavesLdsSys + 1], off_s[kWavesLdsSys];
    const int tid = threadIdx.x;
    if (info && tid < 64) {     // the largest of the update's slots (wave 0), which it clears
        float m = 0.0f;
        if (tid < kStepMaxSlots) {
            m = __uint_as_float(step_max[tid]);
            step_max[tid] = 0u;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
        if (tid == 0) dmax_s[0] = m;
    }
    if (tid == 0) {
        int a = 0;
        if (!lite)      // no list was built: the ranges of the last build stand
            for (int s = 0; s < nsys; ++s) a |= sc[s].rebuild;
        if (a) {
            int t = 0;
            for (int s = 0; s < nsys; ++s) {
                prefix[s] = t;
                if (s < kWavesLdsSys) { pre_s[s] = t; off_s[s] = sd[s].chunk_off; }
                // a system whose build overflowed has an incomplete list (a block that
                // overflowed reserved chunk slots it did not write): no pair work
                if (!sc[s].overflow) t += min(sc[s].n_chunks, sd[s].chunk_cap);
            }
            prefix[nsys] = t;
            if (nsys <= kWavesLdsSys) pre_s[nsys] = t;
            total = t;
        }
        any = a;
    }
    __syncthreads();
    const float dmax2 = (info && tid == 0) ? dmax_s[0] : 0.0f;
    if (pool) waves_pair_begin(sc, nsys, pool, lite, step_seq, info, dmax2);       // the flags were read above
    if (!any) return;
    const int T = total;
    const bool lds = nsys <= kWavesLdsSys;
    auto pre = [&](int s) { return lds ? pre_s[s] : prefix[s]; };
    auto sys_of = [&](int g) {          // largest s with prefix[s] <= g
        int lo = 0, hi = nsys - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (pre(mid) <= g) lo = mid; else hi = mid - 1;
        }
        return lo;
    };
    auto cost_of = [&](int g, int& s) {
        while (g >= pre(s + 1)) ++s;
        return (long long)chunk_cost[(lds ? off_s[s] : sd[s].chunk_off) + g - pre(s)];
    };
    const int per = (T + (int)blockDim.x - 1) / (int)blockDim.x;
    const int g_begin = min(T, tid * per), g_end = min(T, g_begin + per);
    long long mine = 0;
    int s = (g_begin < g_end) ? sys_of(g_begin) : 0;
    for (int g = g_begin; g < g_end; ++g) mine += cost_of(g, s);
    // block exclusive scan of the per-thread sums
    long long incl = mine;
    const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const long long t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    long long base = incl - mine, ctot = 0;
    for (int w2 = 0; w2 < (int)(blockDim.x >> 6); ++w2) {
        if (w2 < wv) base += wsum[w2];
        ctot += wsum[w2];
    }
    if (T == 0) {
        for (int w = tid; w <= NW; w += blockDim.x) wave_start[w] = 0;
    } else if (g_begin < g_end) {
        // doubled units: M_g = 2 * (cost before g) + cost of g; wave of g = M_g * NW / (2 * ctot)
        auto wave_of = [&](long long M) { return (int)min((long long)NW - 1, M * NW / (2 * ctot)); };
        int a_prev = -1;
        if (g_begin > 0) {
            int sp = sys_of(g_begin - 1);
            const long long cp = cost_of(g_begin - 1, sp);
            a_prev = wave_of(2 * (base - cp) + cp);
        }
        long long cum = base;
        s = sys_of(g_begin);
        int a = a_prev;
        for (int g = g_begin; g < g_end; ++g) {
            const long long c = cost_of(g, s);
            a = wave_of(2 * cum + c);
            for (int w = a_prev + 1; w <= a; ++w) wave_start[w] = g;
            a_prev = a;
            cum += c;
        }
        if (g_end == T)
            for (int w = a + 1; w <= NW; ++w) wave_start[w] = T;
    }
}

// Clears the rebuild flags when no pair kernel follows (its first launch does
// it otherwise) and counts the builds.
__global__ void nl_clear_flag(SysCount* sc, int nsys) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    if (sc[s].rebuild) {
        sc[s].rebuild = 0;
        sc[s].rebuilds += 1;
    }
}

template <class P>
void launch_nlist(Batch<P>& B, bool clear_flags, int part) {
    const int T = 256;
    hipStream_t st = B.stream;
    auto nb = [](long n, int t) { return (unsigned)((n + t - 1) / t); };
    if (part == 2 || part == 3) {   // the last kernel alone, doing the pair kernel's set-up too
        // (part 3: in a step without the other list kernels)
        nl_waves<<<1, 1024, 0, st>>>(B.d_sys, B.d_cnt, B.nsys, B.chunk_cost, B.chunk_prefix, B.wave_start, B.pair_waves,
                                     B.pair_pool, part == 3 ? 1 : 0, B.step_seq, B.step_info, B.step_max);
        return;
    }
    nl_bin<<<nb(B.n_atoms, T), T, 0, st>>>(B.d_sys, B.d_cnt, B.atom_sys, B.X, B.Y, B.Z, B.atom_col, B.col_tmp,
                                          B.n_atoms);
    nl_scan<<<B.nsys, 1024, 0, st>>>(B.d_sys, B.d_cnt, B.col_tmp, B.col_count, B.col_slot, B.col_super);
    nl_scatter<<<nb(B.n_atoms, T), T, 0, st>>>(B.d_sys, B.d_cnt, B.atom_sys, B.atom_col, B.col_slot, B.col_fill,
                                              B.slot_tmp, B.n_atoms);
    nl_columns<P><<<B.n_cols, 256, 0, st>>>(B.d_sys, B.d_cnt, B.col_sys, B.col_count, B.col_fill, B.col_slot,
                                            B.col_super, B.slot_tmp, B.X, B.Y, B.Z, B.q, B.sig, B.eps,
                                            B.atom_of_slot, B.slot_of_atom, B.cl_ctr, B.cl_half, B.super_item,
                                            B.slot_zkey, B.slot_wpos, B.pos, B.lj, B.XB, B.YB, B.ZB);
    nl_pairlist<<<B.n_super_cap, kPlThreads, 0, st>>>(B.d_sys, B.d_cnt, B.super_sys, B.super_item, B.col_count, B.col_slot,
                                              B.cl_ctr, B.cl_half, B.slot_wpos, B.atom_of_slot, B.slot_of_atom, B.excl_off,
                                              B.excl_idx, B.list, B.emask, B.chunks, B.chunk_cost, B.n_super_cap);
    if (part == 1) return;  // all but the last kernel
    nl_waves<<<1, 1024, 0, st>>>(B.d_sys, B.d_cnt, B.nsys, B.chunk_cost, B.chunk_prefix, B.wave_start, B.pair_waves,
                                 nullptr, 0, nullptr, nullptr, nullptr);
    if (clear_flags) nl_clear_flag<<<nb(B.nsys, 64), 64, 0, st>>>(B.d_cnt, B.nsys);
}

template void launch_nlist<PrecFP64>(Batch<PrecFP64>&, bool, int);
template void launch_nlist<PrecMixed>(Batch<PrecMixed>&, bool, int);

}  // namespace nx

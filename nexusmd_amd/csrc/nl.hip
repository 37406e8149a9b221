// nl.hip -- neighbour-list build on the device.
//
// Replaces obtainNeighborList / VoxelHash (core/baseline/src/Forces/
// NeighborList.cpp:12-154), which rebuilds a std::map voxel hash every step.
// Here, per rebuild:
//   1. nl_bin      : counting sort of atoms into xy columns (atomic counts);
//   2. nl_scan     : per-system exclusive scan -> slot ranges (columns padded
//                    to 8-atom clusters) and super-cluster ranges;
//   3. nl_scatter  : atoms into their column's slot range;
//   4. nl_columns  : per column, sort by (z, atom id) in LDS -> deterministic
//                    slot order; cluster bounding boxes; super-cluster items;
//   5. nl_pairlist : one wavefront per super-cluster (8 clusters of 8 atoms):
//                    j-clusters of neighbouring columns whose bounding box is
//                    within rlist of an i-cluster (half list: j >= i), an
//                    8-bit mask of interacting i-clusters per entry, and
//                    64-bit exclusion masks for tiles holding excluded pairs.
// The pair kernel applies r^2 < rc^2 exactly per pair, so the evaluated pair
// set is the reference's set whatever the buffer (skin) is.
//
// Every kernel first checks the system's rebuild flag, which the integrator
// raises when an atom has moved more than skin/2 since the last build.
#include "nx_device.h"
#include "nx_engine.h"

namespace nx {

constexpr int kColSortCap = 2048;     // atoms per column sorted in LDS
constexpr int kListCap = 1024;        // pair-list entries per super-cluster (LDS)
constexpr int kExclCap = 2048;        // exclusion triples per super-cluster (LDS)

__global__ void nl_zero(const SysDev* sd, SysCount* sc, int* col_count, int* col_fill, int n_cols_total,
                        const int* col_sys) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cols_total) return;
    int s = col_sys[c];
    if (!sc[s].rebuild) return;
    col_count[c] = 0;
    col_fill[c] = 0;
}

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

__global__ void nl_bin(const SysDev* sd, SysCount* sc, const int* atom_sys, const double* X, const double* Y,
                       const double* Z, int* atom_col, int* col_count, int n_atoms_total) {
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
    atomicAdd(&col_count[col], 1);
}

// One block per system: scan padded column counts.
__global__ void nl_scan(const SysDev* sd, SysCount* sc, const int* col_count, int* col_slot, int* col_super) {
    int s = blockIdx.x;
    if (!sc[s].rebuild) return;
    const SysDev& d = sd[s];
    const int nc = d.ncx * d.ncy;
    __shared__ int tot_slot, tot_sup, carry_slot, carry_sup;
    __shared__ int ws[1024], wsup[1024];
    if (threadIdx.x == 0) { carry_slot = 0; carry_sup = 0; }
    __syncthreads();
    for (int base = 0; base < nc; base += blockDim.x) {
        int c = base + threadIdx.x;
        int cnt = (c < nc) ? col_count[d.col_off + c] : 0;
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
    int k = atomicAdd(&col_fill[col], 1);
    int slot = col_slot[col] + k;
    if (slot < sd[s].slot_off + sd[s].slot_cap) slot_atom_tmp[slot] = a;
}

// One block (256 threads) per column: deterministic z order, cluster boxes, super items.
__global__ __launch_bounds__(256) void nl_columns(const SysDev* sd, SysCount* sc, const int* col_sys,
                                                  const int* col_count, const int* col_slot, const int* col_super,
                                                  const int* slot_atom_tmp, const double* X, const double* Y,
                                                  const double* Z, int* atom_of_slot, int* slot_of_atom,
                                                  float4* cl_ctr, float4* cl_half, int4* super_item,
                                                  float* slot_zkey) {
    const int col = blockIdx.x;
    const int s = col_sys[col];
    if (!sc[s].rebuild || (sc[s].overflow & 1)) return;
    const SysDev& d = sd[s];
    const int cnt = col_count[col];
    const int first = col_slot[col];
    __shared__ double key[kColSortCap];
    __shared__ int ids[kColSortCap];
    if (cnt > kColSortCap) {
        if (threadIdx.x == 0) atomicOr(&sc[s].overflow, 2);
        return;
    }
    int P = 1;
    while (P < cnt) P <<= 1;
    for (int i = threadIdx.x; i < P; i += blockDim.x) {
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
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += blockDim.x) {
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
            atom_of_slot[slot] = ids[i];
            slot_of_atom[ids[i]] = slot;
            // round down so that a key never exceeds the true z (slab binning stays monotone)
            float zk = (float)key[i];
            if ((double)zk > key[i]) zk = nextafterf(zk, -1.0f);
            slot_zkey[slot] = zk;
        } else {
            atom_of_slot[slot] = -1;
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

// One wavefront per super-cluster item.
__global__ __launch_bounds__(64) void nl_pairlist(const SysDev* sd, SysCount* sc, const int* super_sys,
                                                  int4* super_item, const int* col_count, const int* col_slot,
                                                  const float4* cl_ctr, const float4* cl_half,
                                                  const int* atom_of_slot, const int* slot_of_atom,
                                                  const int* excl_off, const int* excl_idx, uint2* list,
                                                  unsigned long long* emask, int4* chunks, int n_items_total) {
    const int item = xcd_remap(blockIdx.x, gridDim.x);
    if (item >= n_items_total) return;
    const int s = super_sys[item];
    if (!sc[s].rebuild || sc[s].overflow) return;
    const SysDev& d = sd[s];
    if (item - d.super_off >= sc[s].n_super) return;
    const int lane = threadIdx.x;
    int4 it = super_item[item];
    const int fc = it.x, ncl = it.y;

    __shared__ unsigned int l_ent[kListCap];      // cj | imask << 24
    __shared__ int l_eid[kListCap];
    __shared__ int n_ent;
    __shared__ unsigned long long x_bits[kExclCap];
    __shared__ int x_entk[kExclCap];
    __shared__ int n_x;
    __shared__ float4 ic[8], ih[8];
    if (lane == 0) { n_ent = 0; n_x = 0; }
    if (lane < 8) {
        if (lane < ncl) { ic[lane] = cl_ctr[fc + lane]; ih[lane] = cl_half[fc + lane]; }
        else { ic[lane] = make_float4(1e30f, 1e30f, 1e30f, 0.f); ih[lane] = make_float4(0.f, 0.f, 0.f, 0.f); }
    }
    __syncthreads();
    const float Lx = (float)d.L[0], Ly = (float)d.L[1], Lz = (float)d.L[2];
    const float rl2 = (float)(d.rlist * d.rlist) * 1.0001f;
    // own column from the first cluster's slot
    const int own_slot = fc * 8;
    int cx0 = 0, cy0 = 0;
    {
        // locate the column through the cluster centre
        float4 c0 = ic[0];
        cx0 = (int)(c0.x / (float)d.colw[0]);
        cy0 = (int)(c0.y / (float)d.colw[1]);
        cx0 = min(max(cx0, 0), d.ncx - 1);
        cy0 = min(max(cy0, 0), d.ncy - 1);
    }
    (void)own_slot;
    const int Dx = (int)ceil((d.rlist + 1e-4) / d.colw[0]) + 1;
    const int Dy = (int)ceil((d.rlist + 1e-4) / d.colw[1]) + 1;
    const int nx_c = (2 * Dx + 1 >= d.ncx) ? d.ncx : 2 * Dx + 1;
    const int ny_c = (2 * Dy + 1 >= d.ncy) ? d.ncy : 2 * Dy + 1;
    const int sx = (2 * Dx + 1 >= d.ncx) ? 0 : cx0 - Dx;
    const int sy = (2 * Dy + 1 >= d.ncy) ? 0 : cy0 - Dy;
    bool overflow = false;
    for (int ix = 0; ix < nx_c; ++ix) {
        int cx = sx + ix;
        cx = ((cx % d.ncx) + d.ncx) % d.ncx;
        for (int iy = 0; iy < ny_c; ++iy) {
            int cy = sy + iy;
            cy = ((cy % d.ncy) + d.ncy) % d.ncy;
            const int col = d.col_off + cx * d.ncy + cy;
            const int cnt = col_count[col];
            const int ccl = (cnt + 7) >> 3;
            const int cfirst = col_slot[col] >> 3;
            for (int b = 0; b < ccl; b += 64) {
                int c = b + lane;
                unsigned int im = 0;
                int cj = cfirst + c;
                if (c < ccl) {
                    float4 jc = cl_ctr[cj], jh = cl_half[cj];
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        if (k < ncl && cj >= fc + k && bbox_gap2(ic[k], ih[k], jc, jh, Lx, Ly, Lz) < rl2) im |= 1u << k;
                    }
                }
                unsigned long long bal = __ballot(im != 0);
                int before = __popcll(bal & ((1ull << lane) - 1ull));
                int tot = __popcll(bal);
                int base = n_ent;
                if (im) {
                    int pos = base + before;
                    if (pos < kListCap) { l_ent[pos] = (unsigned)cj | (im << 24); l_eid[pos] = -1; }
                }
                __syncthreads();
                if (lane == 0) n_ent = base + tot;
                __syncthreads();
            }
        }
    }
    int ne = n_ent;
    if (ne > kListCap) { overflow = true; ne = kListCap; }
    // sort entries by cluster index (bitonic over the next power of two)
    int P = 1;
    while (P < ne) P <<= 1;
    for (int i = ne + lane; i < P; i += 64) { l_ent[i] = 0xffffffffu; l_eid[i] = -1; }
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = lane; i < P; i += 64) {
                int l = i ^ j;
                if (l > i) {
                    bool up = (i & k) == 0;
                    unsigned a = l_ent[i] & 0xffffffu, bb = l_ent[l] & 0xffffffu;
                    if (l_ent[i] == 0xffffffffu) a = 0xffffffffu;
                    if (l_ent[l] == 0xffffffffu) bb = 0xffffffffu;
                    if ((a > bb) == up) { unsigned t = l_ent[i]; l_ent[i] = l_ent[l]; l_ent[l] = t; }
                }
            }
            __syncthreads();
        }
    }
    // exclusions of the 64 i-atoms
    {
        const int k = lane >> 3, ia = lane & 7;
        if (k < ncl) {
            const int si = (fc + k) * 8 + ia;
            const int ai = atom_of_slot[si];
            if (ai >= 0) {
                for (int e = excl_off[ai]; e < excl_off[ai + 1]; ++e) {
                    const int p = excl_idx[e];
                    const int sp = slot_of_atom[p];
                    const int cj = sp >> 3;
                    if (cj < fc + k) continue;
                    if (cj == fc + k && sp <= si) continue;
                    // binary search cj among entries
                    int lo = 0, hi = ne;
                    while (lo < hi) {
                        int m = (lo + hi) >> 1;
                        if ((int)(l_ent[m] & 0xffffffu) < cj) lo = m + 1; else hi = m;
                    }
                    if (lo < ne && (int)(l_ent[lo] & 0xffffffu) == cj && ((l_ent[lo] >> 24) >> k & 1u)) {
                        int t = atomicAdd(&n_x, 1);
                        if (t < kExclCap) {
                            x_bits[t] = 1ull << (ia * 8 + (sp & 7));
                            x_entk[t] = lo * 8 + k;
                        }
                        l_eid[lo] = 0;      // mark
                    }
                }
            }
        }
    }
    __syncthreads();
    int nx_ = n_x;
    if (nx_ > kExclCap) { overflow = true; nx_ = kExclCap; }
    // compact ids for entries with exclusions
    __shared__ int id_base;
    if (lane == 0) {
        int nm = 0;
        for (int e = 0; e < ne; ++e) if (l_eid[e] == 0) l_eid[e] = nm++;
        int b = atomicAdd(&sc[s].emask_count, nm);
        if (b + nm > d.emask_cap) { id_base = -1; }
        else id_base = b;
    }
    __syncthreads();
    if (id_base < 0) overflow = true;
    if (!overflow) {
        for (int e = lane; e < ne; e += 64) {
            if (l_eid[e] >= 0) {
                int g = d.emask_off + id_base + l_eid[e];
                for (int k = 0; k < 8; ++k) emask[(size_t)g * 8 + k] = 0ull;
            }
        }
        __syncthreads();
        for (int t = lane; t < nx_; t += 64) {
            int e = x_entk[t] >> 3, k = x_entk[t] & 7;
            int g = d.emask_off + id_base + l_eid[e];
            atomicOr(&emask[(size_t)g * 8 + k], x_bits[t]);
        }
    }
    __shared__ int lbase, cbase;
    const int nch = (ne + kChunkEntries - 1) / kChunkEntries;
    if (lane == 0) {
        int b = atomicAdd(&sc[s].list_count, ne);
        lbase = (b + ne > d.list_cap) ? -1 : b;
        int c = atomicAdd(&sc[s].n_chunks, nch);
        cbase = (c + nch > d.chunk_cap) ? -1 : c;
    }
    __syncthreads();
    if (lbase < 0 || cbase < 0) overflow = true;
    if (overflow) {
        if (lane == 0) atomicOr(&sc[s].overflow, 4);
        return;
    }
    for (int e = lane; e < ne; e += 64) {
        uint2 v;
        v.x = l_ent[e];
        v.y = (l_eid[e] >= 0) ? (unsigned)(d.emask_off + id_base + l_eid[e]) : 0xffffffffu;
        list[d.list_off + lbase + e] = v;
    }
    if (lane == 0) {
        it.z = d.list_off + lbase;
        it.w = ne;
        super_item[item] = it;
    }
    for (int c = lane; c < nch; c += 64) {
        const int e0 = d.list_off + lbase + c * kChunkEntries;
        const int e1 = d.list_off + lbase + min(ne, (c + 1) * kChunkEntries);
        chunks[d.chunk_off + cbase + c] = make_int4(fc, ncl, e0, e1);
    }
}

// After a build: pack per-slot LJ parameters, record build positions, clear the flag.
template <class P>
__global__ void nl_finish_slots(const SysDev* sd, SysCount* sc, const int* slot_sys, const int* atom_of_slot,
                                const double* sig, const double* eps, typename P::real2* lj, int n_slots_total) {
    int sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl >= n_slots_total) return;
    int s = slot_sys[sl];
    if (!sc[s].rebuild || sc[s].overflow) return;
    int a = (sl - sd[s].slot_off < sc[s].n_slots) ? atom_of_slot[sl] : -1;
    typename P::real2 v;
    if (a >= 0) {
        v.x = (typename P::real)(0.5 * sig[a]);
        v.y = (typename P::real)(2.0 * sqrt(eps[a]));
    } else {
        v.x = 0;
        v.y = 0;
    }
    lj[sl] = v;
}

__global__ void nl_finish_atoms(const SysDev* sd, SysCount* sc, const int* atom_sys, const double* X,
                                const double* Y, const double* Z, double* XB, double* YB, double* ZB,
                                int n_atoms_total) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n_atoms_total) return;
    int s = atom_sys[a];
    if (!sc[s].rebuild || sc[s].overflow) return;
    XB[a] = X[a];
    YB[a] = Y[a];
    ZB[a] = Z[a];
}

__global__ void nl_clear_flag(SysCount* sc, int nsys) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    if (sc[s].rebuild) {
        sc[s].rebuild = 0;
        sc[s].rebuilds += 1;
    }
}

template <class P>
void launch_nlist(Batch<P>& B) {
    const int T = 256;
    hipStream_t st = B.stream;
    auto nb = [](long n, int t) { return (unsigned)((n + t - 1) / t); };
    nl_zero<<<nb(B.n_cols, T), T, 0, st>>>(B.d_sys, B.d_cnt, B.col_count, B.col_fill, B.n_cols, B.col_sys);
    nl_bin<<<nb(B.n_atoms, T), T, 0, st>>>(B.d_sys, B.d_cnt, B.atom_sys, B.X, B.Y, B.Z, B.atom_col, B.col_count,
                                          B.n_atoms);
    nl_scan<<<B.nsys, 1024, 0, st>>>(B.d_sys, B.d_cnt, B.col_count, B.col_slot, B.col_super);
    nl_scatter<<<nb(B.n_atoms, T), T, 0, st>>>(B.d_sys, B.d_cnt, B.atom_sys, B.atom_col, B.col_slot, B.col_fill,
                                              B.slot_tmp, B.n_atoms);
    nl_columns<<<B.n_cols, 256, 0, st>>>(B.d_sys, B.d_cnt, B.col_sys, B.col_count, B.col_slot, B.col_super,
                                         B.slot_tmp, B.X, B.Y, B.Z, B.atom_of_slot, B.slot_of_atom, B.cl_ctr,
                                         B.cl_half, B.super_item, B.slot_zkey);
    nl_pairlist<<<B.n_super_cap, 64, 0, st>>>(B.d_sys, B.d_cnt, B.super_sys, B.super_item, B.col_count, B.col_slot,
                                              B.cl_ctr, B.cl_half, B.atom_of_slot, B.slot_of_atom, B.excl_off,
                                              B.excl_idx, B.list, B.emask, B.chunks, B.n_super_cap);
    nl_finish_slots<P><<<nb(B.n_slots_cap, T), T, 0, st>>>(B.d_sys, B.d_cnt, B.slot_sys, B.atom_of_slot, B.sig,
                                                          B.eps, B.lj, B.n_slots_cap);
    nl_finish_atoms<<<nb(B.n_atoms, T), T, 0, st>>>(B.d_sys, B.d_cnt, B.atom_sys, B.X, B.Y, B.Z, B.XB, B.YB, B.ZB,
                                                   B.n_atoms);
    nl_clear_flag<<<nb(B.nsys, 64), 64, 0, st>>>(B.d_cnt, B.nsys);
}

template void launch_nlist<PrecFP64>(Batch<PrecFP64>&);
template void launch_nlist<PrecMixed>(Batch<PrecMixed>&);

}  // namespace nx

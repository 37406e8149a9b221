// engine.hip -- host orchestration of the device engine and the C-ABI.
//
// Replaces BaseLine::Engine (core/baseline/src/Engine.cpp:19-188): topology
// and box set-up once, then per step  [neighbour-list rebuild if flagged] ->
// pack -> bonded -> exclusion correction -> pair -> PME -> (report-step
// kinetic energy) -> leapfrog, all enqueued on one HIP stream; the
// report-free step is captured once into a hipGraph and replayed.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nexusmd.h"
#include <cstdlib>

#include "nx_device.h"
#include "nx_engine.h"
#include "nx_host.h"

namespace nx {

#define HIPCHK(x)                                                                                        \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess)                                                                            \
            throw ::nx::HipError(e_ == hipErrorOutOfMemory ? NX_ERR_NOMEM : NX_ERR_HIP,                  \
                                 std::string(#x) + ": " + hipGetErrorString(e_));                        \
    } while (0)

template <class T>
static void dalloc(T** p, size_t n, std::vector<void*>& owned, hipStream_t st) {
    if (n == 0) n = 1;
    HIPCHK(hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T)));
    owned.push_back(*p);
    // zeroed, so capacity-overflow gaps in work arrays read as empty items; on the
    // engine's stream, which does not order with the null stream, so that the
    // clear lands before the uploads and kernels queued after it
    HIPCHK(hipMemsetAsync(*p, 0, n * sizeof(T), st));
}

template <class T>
static void upload(T* d, const std::vector<T>& h, hipStream_t st) {
    if (!h.empty()) HIPCHK(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st));
}

// ------------------------------------------------------------- host setup

// 1-2, 1-3 and 1-4 partners (graph distance <= 3 over the bond graph), the
// set Exclusions::createExclusions builds with bondCutoff = 3
// (core/baseline/src/Forces/Exclusions.cpp:7-44).
void exclusions_from_bonds(int n, const std::vector<nx_bond>& bonds, std::vector<int>& off, std::vector<int>& idx) {
    std::vector<std::vector<int>> adj(n);
    for (const auto& b : bonds) {
        if (b.p1 < 0 || b.p2 < 0 || b.p1 >= n || b.p2 >= n)
            throw HipError(NX_ERR_ARG, "createExclusions: Illegal particle index in list of bonds");
        adj[b.p1].push_back(b.p2);
        adj[b.p2].push_back(b.p1);
    }
    off.assign(n + 1, 0);
    idx.clear();
    std::vector<int> mark(n, -1), frontier, nxt, found;
    for (int i = 0; i < n; ++i) {
        found.clear();
        frontier.assign(1, i);
        mark[i] = i;
        for (int depth = 0; depth < 3; ++depth) {
            nxt.clear();
            for (int u : frontier)
                for (int v : adj[u])
                    if (mark[v] != i) {
                        mark[v] = i;
                        nxt.push_back(v);
                        found.push_back(v);
                    }
            frontier.swap(nxt);
        }
        std::sort(found.begin(), found.end());
        idx.insert(idx.end(), found.begin(), found.end());
        off[i + 1] = (int)idx.size();
    }
}

// |b(m)|^2 of the order-5 B-spline by the O(K^2) DFT and neighbour-average
// fix-up of NonbondedForce::computeBSplineParameters (NonbondedForce.cpp:77-142).
static void bspline_moduli(int K, std::vector<double>& mod) {
    const int order = 5;
    double data[5] = {0, 0, 0, 0, 0};
    data[order - 1] = 0.0;
    data[1] = 0.0;
    data[0] = 1.0;
    for (int i = 3; i < order; ++i) {
        double div = 1.0 / (i - 1);
        data[i - 1] = 0.0;
        for (int j = 1; j < (i - 1); ++j) data[i - j - 1] = div * (j * data[i - j - 2] + (i - j) * data[i - j - 1]);
        data[0] = div * data[0];
    }
    double div = 1.0 / (order - 1);
    data[order - 1] = div * data[order - 2];
    for (int i = 1; i < (order - 1); ++i)
        data[order - i - 1] = div * (i * data[order - i - 2] + (order - i) * data[order - i - 1]);
    data[0] = div * data[0];
    std::vector<double> bd(std::max(K, order + 1), 0.0);
    for (int i = 1; i <= order; ++i) bd[i] = data[i - 1];
    mod.assign(K, 0.0);
    for (int i = 0; i < K; ++i) {
        double sc = 0.0, ss = 0.0;
        for (int j = 0; j < K; ++j) {
            double arg = (2 * kPi * i * j) / K;
            sc += bd[j] * cos(arg);
            ss += bd[j] * sin(arg);
        }
        mod[i] = sc * sc + ss * ss;
    }
    for (int i = 0; i < K; ++i)
        if (mod[i] < 1e-7) mod[i] = (mod[(i - 1 + K) % K] + mod[(i + 1) % K]) / 2.0;
}

static bool fft_size_ok(int K) {
    if (K < 2) return false;
    int m = K;
    for (int p : {2, 3, 5})
        while (m % p == 0) m /= p;
    return m == 1;
}

// PME grid rule: next power of two >= 10 * ceil(L) per axis (NonbondedForce.cpp:33-48).
static int grid_rule(double L) {
    int m = (int)ceil(L) * 10, p = 1;
    while (p < m) p *= 2;
    return p;
}

// ----------------------------------------------------------- tuning cache
//
// Launch shapes chosen by autotune, by key, for the process; with
// NEXUSMD_TUNE_CACHE naming a file, also read from and appended to that file
// (one "key<TAB>pair spread xblocks lines" line per entry, the last one of a
// key winning), so that later processes start from them.
namespace {
std::mutex g_tune_mu;
std::map<std::string, Tuning> g_tune;
bool g_tune_loaded = false;

void tune_load_locked() {
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    const char* path = std::getenv("NEXUSMD_TUNE_CACHE");
    if (!path || !*path) return;
    FILE* f = std::fopen(path, "r");
    if (!f) return;
    char line[1024];
    while (std::fgets(line, sizeof(line), f)) {
        char* tab = std::strchr(line, '\t');
        if (!tab) continue;
        *tab = 0;
        Tuning t;
        if (std::sscanf(tab + 1, "%d %d %d %d", &t.pair_per_cu, &t.spread_block, &t.fft_xblocks, &t.fft_lines) == 4)
            g_tune[line] = t;
    }
    std::fclose(f);
}
}  // namespace

bool tune_cache_get(const std::string& key, Tuning* t) {
    if (key.empty()) return false;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_load_locked();
    auto it = g_tune.find(key);
    if (it == g_tune.end()) return false;
    *t = it->second;
    return true;
}

void tune_cache_put(const std::string& key, const Tuning& t) {
    if (key.empty()) return;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_load_locked();
    g_tune[key] = t;
    const char* path = std::getenv("NEXUSMD_TUNE_CACHE");
    if (!path || !*path) return;
    if (FILE* f = std::fopen(path, "a")) {
        std::fprintf(f, "%s\t%d %d %d %d\n", key.c_str(), t.pair_per_cu, t.spread_block, t.fft_xblocks, t.fft_lines);
        std::fclose(f);
    }
}

// ----------------------------------------------------------------- engine

template <class P>
class Engine : public EngineBase {
   public:
    using real = typename P::real;
    using real2 = typename P::real2;
    using real4 = typename P::real4;
    using facc = typename P::facc;

    Engine(int device) : device_(device) {
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipStreamCreateWithFlags(&B.stream, hipStreamNonBlocking));
        // NEXUSMD_GRAPHS=0: every step as direct launches (no graph capture or replay)
        const char* g = std::getenv("NEXUSMD_GRAPHS");
        if (g && std::string(g) == "0") graphs_enabled_ = false;
        // PME runs on a second stream beside the pair kernel; NEXUSMD_OVERLAP=0 serialises them
        const char* o = std::getenv("NEXUSMD_OVERLAP");
        if (o && std::string(o) == "0") overlap_ = false;
        // NEXUSMD_PAIR_IMAGE=0: the mixed pair kernel takes every pair's nearest
        // image itself instead of the list's per-entry image
        const char* pi = std::getenv("NEXUSMD_PAIR_IMAGE");
        if (pi && std::string(pi) == "0") pair_image_ = false;
        HIPCHK(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&join_, hipEventDisableTiming));
        // NEXUSMD_LITE=0: every step graph carries the list kernels; =2: the step
        // without them is chosen without the forecast (every build then goes
        // through a frozen step and its repeat; for tests of that path)
        const char* li = std::getenv("NEXUSMD_LITE");
        if (li && std::string(li) == "0") lite_mode_ = 0;
        if (li && std::string(li) == "2") lite_mode_ = 2;
    }
    ~Engine() override {
        hipSetDevice(device_);
        release();
        if (B.stream) hipStreamDestroy(B.stream);
        if (side_) hipStreamDestroy(side_);
        if (fork_) hipEventDestroy(fork_);
        if (join_) hipEventDestroy(join_);
        if (info_host_) hipHostFree(info_host_);
        for (auto& e : evpool_) hipEventDestroy(e);
        for (auto& e : step_ev_) hipEventDestroy(e);
        if (copy_) hipStreamDestroy(copy_);
        if (packed_) hipEventDestroy(packed_);
    }

    void release() {
        free_reports();
        destroy_graphs();
        destroy_segments();
        for (void* p : owned_) hipFree(p);
        owned_.clear();
        Batch<P> fresh;
        fresh.stream = B.stream;
        fresh.timing = B.timing;
        B = fresh;
    }

    // Lay out and upload a batch of systems.
    void setup(const std::vector<HostSystem*>& systems) override {
        hipSetDevice(device_);
        release();
        hs_ = systems;
        if (!tuning_set_) {
            // launch shapes an earlier autotune chose for the same kind of system
            Tuning t;
            tuning_ = tune_cache_get(tune_key(), &t) ? t : Tuning{};
        }
        const int ns = (int)systems.size();
        B.nsys = ns;
        HIPCHK(hipDeviceGetAttribute(&B.n_cu, hipDeviceAttributeMultiprocessorCount, device_));
        B.sys.assign(ns, SysDev{});
        long long atom_off = 0, slot_off = 0, col_off = 0, sup_off = 0, list_off = 0, emask_off = 0, chunk_off = 0;
        long long rg_off = 0, cg_off = 0, tw_off = 0, excl_off_total = 0;
        int bond_off = 0, angle_off = 0, tors_off = 0;
        fft_maxK_ = 0;
        for (int s = 0; s < ns; ++s) {
            HostSystem& h = *systems[s];
            SysDev& d = B.sys[s];
            const int n = h.n;
            for (int k = 0; k < 3; ++k) {
                d.lb[k] = h.lb[k];
                d.ub[k] = h.ub[k];
                d.L[k] = fabs(h.ub[k] - h.lb[k]);
                d.invL[k] = 1 / d.L[k];
            }
            const double rc = h.nb.cutoff;
            d.rc = rc;
            d.rc2 = rc * rc;
            const double skin = std::max(0.0, h.nb.skin);
            d.rlist = rc + skin;
            d.skin_half2 = skin > 0 ? 0.25 * skin * skin : -1.0;
            d.alpha = (1.0 / rc) * sqrt(-log(2.0 * h.nb.ewald_tolerance));
            if (h.nb.use_switch && h.nb.switch_distance < rc) {
                d.sw_on = h.nb.switch_distance;
                d.sw_inv = 1.0 / (rc - h.nb.switch_distance);
            } else {
                d.sw_on = -1.0;
                d.sw_inv = 0.0;
            }
            d.n_atoms = n;
            d.atom_off = (int)atom_off;
            d.pair_image = pair_image_ ? 1 : 0;
            // columns of ~8/rho^(1/3) nm so that a cluster of 8 is roughly cubic
            const double V = d.L[0] * d.L[1] * d.L[2];
            const double rho = std::max(1e-3, n / V);
            double cw = cbrt(8.0 / rho);
            int ncx = std::max(1, (int)floor(d.L[0] / cw));
            int ncy = std::max(1, (int)floor(d.L[1] / cw));
            while ((double)n / ((double)ncx * ncy) > 700.0) { ncx += 1; ncy += 1; }
            d.ncx = ncx;
            d.ncy = ncy;
            d.colw[0] = d.L[0] / ncx;
            d.colw[1] = d.L[1] / ncy;
            d.col_off = (int)col_off;
            d.col_cap = ncx * ncy;
            d.slot_off = (int)slot_off;
            d.slot_cap = ((n + 8 * ncx * ncy + 63) / 64) * 64;
            d.super_off = (int)sup_off;
            d.super_cap = (n + 63) / 64 + ncx * ncy;
            // pair-list capacity from density: clusters within the column neighbourhood
            const double w = std::max(d.colw[0], d.colw[1]);
            const double hz = std::min(d.L[2], 64.0 / (rho * d.colw[0] * d.colw[1]));
            const double r = d.rlist;
            double ext = (std::min(d.L[0], w + 2 * r + 2 * w)) * (std::min(d.L[1], w + 2 * r + 2 * w)) *
                         std::min(d.L[2], hz + 2 * r + 0.5);
            double per_super = 1.5 * (rho / 8.0) * ext * h.list_scale;
            per_super = std::min(per_super, (double)kSuperListCap);
            per_super = std::max(per_super, 16.0);
            d.list_off = (int)list_off;
            d.list_cap = (int)std::min<double>(2.0e9, per_super * d.super_cap + 64);
            d.chunk_off = (int)chunk_off;
            d.chunk_cap = d.list_cap / kChunkEntries + d.super_cap;
            d.emask_off = (int)emask_off;
            d.emask_cap = (int)(std::max<long long>(256, (long long)h.excl_idx.size() / 2 + 64) * h.list_scale);
            // PME grid
            int g[3];
            for (int k = 0; k < 3; ++k) g[k] = h.nb.grid[k] > 0 ? h.nb.grid[k] : grid_rule(d.L[k]);
            for (int k = 0; k < 3; ++k)
                if (!fft_size_ok(g[k]) || g[k] < 6 || g[k] > 1024)
                    throw HipError(NX_ERR_ARG, "PME grid dimension " + std::to_string(g[k]) +
                                                   " must be in [6, 1024] with prime factors 2, 3, 5 only");
            for (int k = 0; k < 3; ++k) d.grid[k] = g[k];
            d.rgrid_off = rg_off;
            d.cgrid_off = cg_off;
            for (int k = 0; k < 3; ++k) {
                d.tw_off[k] = (int)tw_off;
                tw_off += g[k];
                fft_maxK_ = std::max(fft_maxK_, g[k]);
            }
            // spread tiling: column blocks b x b and nz slabs whose LDS region fits kSpreadLdsPoints
            {
                d.sp_margin = 0.5 * skin + 0.002;
                long long best = -1;
                int bb = 0, bnz = 0;
                const int bmax = std::min(12, std::max(ncx, ncy));
                // a tuned block size is taken when it fits; otherwise the cost model chooses
                const int want = tuning_.spread_block > 0 ? std::min(tuning_.spread_block, bmax) : 0;
                for (int pass = want > 0 ? 0 : 1; pass < 2 && best < 0; ++pass)
                for (int b = pass == 0 ? want : 1; b <= (pass == 0 ? want : bmax); ++b) {
                    for (int nz = 1; nz <= g[2]; ++nz) {
                        auto ext = [&](double w, int K, double Lk, bool full) {
                            if (full) return K;
                            const double hh = Lk / K;
                            int n = (int)ceil((w + 2 * d.sp_margin) / hh) + 6;
                            return n >= K ? K : n;
                        };
                        const int nx = ext(b * d.colw[0], g[0], d.L[0], b >= ncx);
                        const int ny = ext(b * d.colw[1], g[1], d.L[1], b >= ncy);
                        const int nzp = ext(d.L[2] / nz, g[2], d.L[2], nz == 1);
                        const long long pts = (long long)nx * spread_plane(ny, nzp);
                        if (pts > kSpreadLdsPoints) continue;
                        const long long tiles = (long long)((ncx + b - 1) / b) * ((ncy + b - 1) / b) * nz;
                        const long long cost = pts * tiles + 2000 * (tiles < 128 ? 128 - tiles : 0);
                        if (best < 0 || cost < best) { best = cost; bb = b; bnz = nz; }
                        break;   // the smallest fitting nz for this b
                    }
                }
                d.sp_block[0] = bb;
                d.sp_block[1] = bb;
                d.sp_block[2] = bnz;
            }
            rg_off += (long long)g[0] * g[1] * g[2];
            cg_off += (long long)g[0] * g[1] * (g[2] / 2 + 1);
            d.bond_off = bond_off;
            d.n_bonds = (int)h.bonds.size();
            d.angle_off = angle_off;
            d.n_angles = (int)h.angles.size();
            d.tors_off = tors_off;
            d.n_tors = (int)h.torsions.size();
            bond_off += d.n_bonds;
            angle_off += d.n_angles;
            tors_off += d.n_tors;
            atom_off += n;
            slot_off += d.slot_cap;
            col_off += (long long)ncx * ncy;
            sup_off += d.super_cap;
            list_off += d.list_cap;
            chunk_off += d.chunk_cap;
            emask_off += d.emask_cap;
            excl_off_total += (long long)h.excl_idx.size();
        }
        // cluster ids fill 24 bits of a list word, exclusion-mask ids 26
        if (slot_off > (1 << 27) || list_off > 2000000000LL || emask_off >= (long long)kNoMask)
            throw HipError(NX_ERR_ARG, "batch too large");
        B.n_atoms = (int)atom_off;
        B.n_slots_cap = (int)slot_off;
        B.n_cols = (int)col_off;
        B.n_super_cap = (int)sup_off;
        B.n_chunk_cap = (int)chunk_off;
        B.n_list_cap = list_off;
        B.n_emask_cap = emask_off;
        B.rgrid_total = rg_off;
        B.cgrid_total = cg_off;
        B.tw_total = tw_off;
        B.n_bonds = bond_off;
        B.n_angles = angle_off;
        B.n_tors = tors_off;
        B.pme_any = 1;
        allocate();
        upload_all();
    }

    void allocate() {
        auto& o = owned_;
        dalloc(&B.d_sys, B.nsys, o, B.stream);
        dalloc(&B.d_cnt, B.nsys, o, B.stream);
        const size_t A = B.n_atoms, S = B.n_slots_cap;
        for (double** p : {&B.X, &B.Y, &B.Z, &B.VX, &B.VY, &B.VZ, &B.invm, &B.mass, &B.q, &B.sig, &B.eps, &B.XB, &B.YB,
                           &B.ZB})
            dalloc(p, A, o, B.stream);
        dalloc(&B.atom_sys, A, o, B.stream);
        dalloc(&B.slot_of_atom, A, o, B.stream);
        dalloc(&B.atom_col, A, o, B.stream);
        dalloc(&B.FAx, A, o, B.stream);
        dalloc(&B.FAy, A, o, B.stream);
        dalloc(&B.FAz, A, o, B.stream);
        dalloc(&B.slot_sys, S, o, B.stream);
        dalloc(&B.atom_of_slot, S, o, B.stream);
        dalloc(&B.slot_tmp, S, o, B.stream);
        dalloc(&B.pos, S, o, B.stream);
        dalloc(&B.lj, S, o, B.stream);
        dalloc(&B.FSx, S, o, B.stream);
        dalloc(&B.FSy, S, o, B.stream);
        dalloc(&B.FSz, S, o, B.stream);
        for (int** p : {&B.col_sys, &B.col_count, &B.col_fill, &B.col_slot, &B.col_super, &B.col_tmp})
            dalloc(p, B.n_cols, o, B.stream);
        dalloc(&B.cl_ctr, S / 8 + 1, o, B.stream);
        dalloc(&B.cl_half, S / 8 + 1, o, B.stream);
        dalloc(&B.super_sys, B.n_super_cap, o, B.stream);
        dalloc(&B.super_item, B.n_super_cap, o, B.stream);
        dalloc(&B.chunk_sys, B.n_chunk_cap, o, B.stream);
        dalloc(&B.slot_zkey, S, o, B.stream);
        dalloc(&B.slot_wpos, S, o, B.stream);
        dalloc(&B.chunks, B.n_chunk_cap, o, B.stream);
        dalloc(&B.chunk_prefix, B.nsys + 1, o, B.stream);
        dalloc(&B.chunk_cost, B.n_chunk_cap, o, B.stream);
        B.pair_blocks = pair_grid_blocks<P>(B.n_cu, B.n_chunk_cap, tuning_.pair_per_cu);
        B.pair_waves = B.pair_blocks * kPairWaves;
        dalloc(&B.wave_start, (size_t)B.pair_waves + 1, o, B.stream);
        dalloc(&B.pair_pool, (size_t)kPoolInts, o, B.stream);
        if (B.nsys == 1 && lite_mode_) {
            // the update's report to the host, for the choice of step graph
            if (!info_host_) {
                HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&info_host_), sizeof(StepInfo),
                                     hipHostMallocMapped | hipHostMallocCoherent));
            }
            std::memset(info_host_, 0, sizeof(StepInfo));
            HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void**>(&B.step_info), info_host_, 0));
            dalloc(&B.step_seq, 1, o, B.stream);
            dalloc(&B.step_max, (size_t)kStepMaxSlots, o, B.stream);
            grow_n_ = 0;
            last_seq_ = 0;
        }
        force_full_ = true;
        dalloc(&B.list, (size_t)B.n_list_cap, o, B.stream);
        dalloc(&B.emask, (size_t)B.n_emask_cap * 8, o, B.stream);
        dalloc(&B.rgrid, (size_t)B.rgrid_total, o, B.stream);
        dalloc(&B.cgrid, (size_t)B.cgrid_total, o, B.stream);
        dalloc(&B.phi, (size_t)B.rgrid_total, o, B.stream);
        dalloc(&B.eta, (size_t)B.cgrid_total, o, B.stream);
        dalloc(&B.twiddle, (size_t)B.tw_total, o, B.stream);
        dalloc(&B.energy, (size_t)B.nsys * E_NTERMS, o, B.stream);
        dalloc(&d_count_, 1, o, B.stream);
        if (B.n_bonds) { dalloc(&B.bond_idx, B.n_bonds, o, B.stream); dalloc(&B.bond_par, B.n_bonds, o, B.stream); }
        if (B.n_angles) { dalloc(&B.angle_idx, B.n_angles, o, B.stream); dalloc(&B.angle_par, B.n_angles, o, B.stream); }
        if (B.n_tors) { dalloc(&B.tors_idx, B.n_tors, o, B.stream); dalloc(&B.tors_par, B.n_tors, o, B.stream); }
        HIPCHK(hipDeviceSynchronize());
    }

    void upload_all() {
        hipStream_t st = B.stream;
        const int ns = B.nsys;
        std::vector<double> mass, invm, q, sig, eps;
        std::vector<int> atom_sys, slot_sys, col_sys, super_sys, chunk_sys, eoff(1, 0), eidx;
        std::vector<int2> epairs, bidx;
        std::vector<double2> bpar, apar;
        std::vector<int4> aidx, tidx;
        std::vector<double4> tpar;
        std::vector<real> eta;
        std::vector<real2> tw;
        std::vector<int4> zi, yi, xi, sp_tiles;
        int direct_slots = 0;
        const int maxK = fft_maxK_;
        // LDS per FFT block: two buffers of T lines plus the twiddles, within 60 KiB
        const int Tfit = std::max(1, std::min(8, (int)((61440 / (int)sizeof(real2) - maxK) / (2 * maxK))));
        const int T = tuning_.fft_lines > 0 ? std::min(tuning_.fft_lines, Tfit) : Tfit;
        B.fft_lines = T;
        B.fft_tile = T;
        B.fft_maxK = maxK;
        // plane passes when every system's planes fit one workgroup's LDS; the x
        // pass splits the kz range so that the batch has about 512 x blocks
        bool plane = fft_plane_enabled();
        int sum_ky = 0;
        for (int s = 0; s < ns; ++s) {
            const SysDev& d = B.sys[s];
            if (plane_lds_bytes<real2>(d.grid[1], d.grid[2]) > kMaxBlockLds ||
                xpass_lds_bytes<real2>(d.grid[0], 1) > kMaxBlockLds)
                plane = false;
            sum_ky += d.grid[1];
        }
        const int xtarget = tuning_.fft_xblocks > 0 ? tuning_.fft_xblocks : 512;
        const int xsplit = std::max(1, (xtarget + sum_ky - 1) / std::max(1, sum_ky));
        B.fft_plane = plane ? 1 : 0;
        B.fft_plane_lds = 0;
        B.fft_x_lds = 0;
        for (int s = 0; s < ns; ++s) {
            HostSystem& h = *hs_[s];
            SysDev& d = B.sys[s];
            const int a0 = d.atom_off;
            const int excl_begin = (int)epairs.size();
            for (int i = 0; i < h.n; ++i) {
                mass.push_back(h.mass[i]);
                invm.push_back(h.mass[i] == 0.0 ? 0.0 : 1.0 / h.mass[i]);
                q.push_back(h.q[i]);
                sig.push_back(h.sig[i]);
                eps.push_back(h.eps[i]);
                atom_sys.push_back(s);
                for (int e = h.excl_off[i]; e < h.excl_off[i + 1]; ++e) {
                    int j = h.excl_idx[e];
                    eidx.push_back(a0 + j);
                    if (j > i) epairs.push_back(make_int2(a0 + i, a0 + j));
                }
                eoff.push_back((int)eidx.size());
            }
            d.excl_off = excl_begin;
            d.n_excl = (int)epairs.size() - excl_begin;
            for (int i = 0; i < d.slot_cap; ++i) slot_sys.push_back(s);
            for (int i = 0; i < d.ncx * d.ncy; ++i) col_sys.push_back(s);
            for (int i = 0; i < d.super_cap; ++i) super_sys.push_back(s);
            for (int i = 0; i < d.chunk_cap; ++i) chunk_sys.push_back(s);
            for (auto& b : h.bonds) { bidx.push_back(make_int2(a0 + b.p1, a0 + b.p2)); bpar.push_back(make_double2(b.d, b.k)); }
            for (auto& a : h.angles) {
                aidx.push_back(make_int4(a0 + a.p1, a0 + a.p2, a0 + a.p3, 0));
                apar.push_back(make_double2(a.a, a.k));
            }
            for (auto& t : h.torsions) {
                tidx.push_back(make_int4(a0 + t.p1, a0 + t.p2, a0 + t.p3, a0 + t.p4));
                tpar.push_back(make_double4(t.k, t.phase, (double)t.periodicity, 0.0));
            }
            // PME tables
            const int Kx = d.grid[0], Ky = d.grid[1], Kz = d.grid[2], Kzh = Kz / 2 + 1;
            std::vector<double> mod[3];
            bspline_moduli(Kx, mod[0]);
            bspline_moduli(Ky, mod[1]);
            bspline_moduli(Kz, mod[2]);
            const double factor = kPi * kPi / (d.alpha * d.alpha);
            const double boxfactor = kPi * d.L[0] * d.L[1] * d.L[2];
            const int mkx = (Kx + 1) / 2, mky = (Ky + 1) / 2, mkz = (Kz + 1) / 2;
            for (int kx = 0; kx < Kx; ++kx) {
                const double mx = (kx < mkx) ? kx : (kx - Kx);
                const double mhx = mx * d.invL[0];
                const double bx = boxfactor * mod[0][kx];
                for (int ky = 0; ky < Ky; ++ky) {
                    const double my = (ky < mky) ? ky : (ky - Ky);
                    const double mhy = my * d.invL[1];
                    const double by = mod[1][ky];
                    for (int kz = 0; kz < Kzh; ++kz) {
                        if (kx == 0 && ky == 0 && kz == 0) { eta.push_back((real)0); continue; }
                        const double mz = (kz < mkz) ? kz : (kz - Kz);
                        const double mhz = mz * d.invL[2];
                        const double m2 = mhx * mhx + mhy * mhy + mhz * mhz;
                        const double bz = mod[2][kz];
                        const double denom = m2 * bx * by * bz;
                        eta.push_back((real)(kOne4PiEps0 * exp(-factor * m2) / denom));
                    }
                }
            }
            const int K3[3] = {Kx, Ky, Kz};
            for (int a = 0; a < 3; ++a)
                for (int t = 0; t < K3[a]; ++t) {
                    const double ang = -2.0 * kPi * (double)t / (double)K3[a];
                    real2 v;
                    v.x = (real)cos(ang);
                    v.y = (real)sin(ang);
                    tw.push_back(v);
                }
            if (d.sp_block[0] > 0) {
                const int b = d.sp_block[0], nz = d.sp_block[2];
                for (int bx = 0; bx < (d.ncx + b - 1) / b; ++bx)
                    for (int by = 0; by < (d.ncy + b - 1) / b; ++by)
                        for (int bz = 0; bz < nz; ++bz) sp_tiles.push_back(make_int4(s, bx, by, bz));
            } else {
                direct_slots += d.slot_cap;
            }
            if (B.fft_plane) {
                for (int x = 0; x < Kx; ++x) zi.push_back(make_int4(s, x, 0, 0));
                int nk = (Kzh + xsplit - 1) / xsplit;
                while (nk > 1 && xpass_lds_bytes<real2>(Kx, nk) > kMaxBlockLds) nk = (nk + 1) / 2;
                for (int y = 0; y < Ky; ++y)
                    for (int k = 0; k < Kzh; k += nk) xi.push_back(make_int4(s, y, k, std::min(nk, Kzh - k)));
                B.fft_plane_lds = std::max(B.fft_plane_lds, plane_lds_bytes<real2>(Ky, Kz));
                B.fft_x_lds = std::max(B.fft_x_lds, xpass_lds_bytes<real2>(Kx, nk));
            } else {
                for (int l = 0; l < Kx * Ky; l += T) zi.push_back(make_int4(s, l, std::min(T, Kx * Ky - l), 0));
                for (int x = 0; x < Kx; ++x)
                    for (int k = 0; k < Kzh; k += T) yi.push_back(make_int4(s, x, k, 0));
                for (int y = 0; y < Ky; ++y)
                    for (int k = 0; k < Kzh; k += T) xi.push_back(make_int4(s, y, k, 0));
            }
            self_energy_.resize(ns);
            double se = 0.0;
            for (int i = 0; i < h.n; ++i) se -= kOne4PiEps0 * h.q[i] * h.q[i] * d.alpha / kSqrtPi;
            self_energy_[s] = se;
        }
        // term-force slots: bonds 2, angles 3, torsions 4 per term, in that order;
        // per atom the slots acting on it (CSR, ascending slot order)
        {
            const int nb = (int)bidx.size(), na = (int)aidx.size(), nt = (int)tidx.size();
            B.n_tslots = 2 * nb + 3 * na + 4 * nt;
            std::vector<int> cnt(B.n_atoms + 1, 0);
            auto each = [&](auto&& fn) {
                for (int t = 0; t < nb; ++t) { fn(bidx[t].x, 2 * t); fn(bidx[t].y, 2 * t + 1); }
                for (int t = 0; t < na; ++t) {
                    const int b = 2 * nb + 3 * t;
                    fn(aidx[t].x, b); fn(aidx[t].y, b + 1); fn(aidx[t].z, b + 2);
                }
                for (int t = 0; t < nt; ++t) {
                    const int b = 2 * nb + 3 * na + 4 * t;
                    fn(tidx[t].x, b); fn(tidx[t].y, b + 1); fn(tidx[t].z, b + 2); fn(tidx[t].w, b + 3);
                }
            };
            each([&](int a, int) { ++cnt[a + 1]; });
            for (int a = 0; a < B.n_atoms; ++a) cnt[a + 1] += cnt[a];
            std::vector<int> at_idx(std::max(1, B.n_tslots)), fill(cnt.begin(), cnt.end() - 1);
            each([&](int a, int slot) { at_idx[fill[a]++] = slot; });
            dalloc(&B.at_off, cnt.size(), owned_, B.stream);
            dalloc(&B.at_idx, at_idx.size(), owned_, B.stream);
            dalloc(&B.tF, (size_t)3 * std::max(1, B.n_tslots), owned_, B.stream);
            upload(B.at_off, cnt, st);
            upload(B.at_idx, at_idx, st);
        }
        B.n_excl_pairs = (int)epairs.size();
        dalloc(&B.excl_off, eoff.size(), owned_, B.stream);
        dalloc(&B.excl_idx, std::max<size_t>(1, eidx.size()), owned_, B.stream);
        if (B.n_excl_pairs) dalloc(&B.excl_pairs, epairs.size(), owned_, B.stream);
        B.n_spread_tiles = (int)sp_tiles.size();
        B.n_spread_direct_slots = direct_slots;
        dalloc(&B.spread_tiles, std::max<size_t>(1, sp_tiles.size()), owned_, B.stream);
        upload(B.spread_tiles, sp_tiles, st);
        dalloc(&B.fftz_items, zi.size(), owned_, B.stream);
        dalloc(&B.ffty_items, yi.size(), owned_, B.stream);
        dalloc(&B.fftx_items, xi.size(), owned_, B.stream);
        B.n_fftz = (int)zi.size();
        B.n_ffty = (int)yi.size();
        B.n_fftx = (int)xi.size();
        HIPCHK(prepare_fft<P>(B));
        upload(B.d_sys, B.sys, st);
        upload(B.mass, mass, st);
        upload(B.invm, invm, st);
        upload(B.q, q, st);
        upload(B.sig, sig, st);
        upload(B.eps, eps, st);
        upload(B.atom_sys, atom_sys, st);
        upload(B.slot_sys, slot_sys, st);
        upload(B.col_sys, col_sys, st);
        upload(B.super_sys, super_sys, st);
        upload(B.chunk_sys, chunk_sys, st);
        upload(B.excl_off, eoff, st);
        upload(B.excl_idx, eidx, st);
        upload(B.excl_pairs, epairs, st);
        upload(B.bond_idx, bidx, st);
        upload(B.bond_par, bpar, st);
        upload(B.angle_idx, aidx, st);
        upload(B.angle_par, apar, st);
        upload(B.tors_idx, tidx, st);
        upload(B.tors_par, tpar, st);
        upload(B.eta, eta, st);
        upload(B.twiddle, tw, st);
        upload(B.fftz_items, zi, st);
        upload(B.ffty_items, yi, st);
        upload(B.fftx_items, xi, st);
        HIPCHK(hipMemsetAsync(B.d_cnt, 0, sizeof(SysCount) * B.nsys, st));
        HIPCHK(hipMemsetAsync(B.FAx, 0, sizeof(facc) * B.n_atoms, st));
        HIPCHK(hipMemsetAsync(B.FAy, 0, sizeof(facc) * B.n_atoms, st));
        HIPCHK(hipMemsetAsync(B.FAz, 0, sizeof(facc) * B.n_atoms, st));
        HIPCHK(hipMemsetAsync(B.atom_of_slot, 0xff, sizeof(int) * B.n_slots_cap, st));
        HIPCHK(hipMemsetAsync(B.slot_of_atom, 0, sizeof(int) * B.n_atoms, st));
        HIPCHK(hipMemsetAsync(B.super_item, 0, sizeof(int4) * B.n_super_cap, st));
        // positions and velocities
        // The first build runs on each system's carried-over build positions
        // when it has them and no rebuild was due, else on its current positions.
        std::vector<double> x, y, z, vx, vy, vz, bx, by, bz;
        bool carried = false;
        for (int s = 0; s < ns; ++s) {
            HostSystem& h = *hs_[s];
            const bool use_b = !h.build_pos.empty() && !h.rebuild_pending;
            const double* b = use_b ? h.build_pos.data() : h.pos.data();
            carried |= use_b;
            for (int i = 0; i < h.n; ++i) {
                x.push_back(h.pos[3 * i]); y.push_back(h.pos[3 * i + 1]); z.push_back(h.pos[3 * i + 2]);
                vx.push_back(h.vel[3 * i]); vy.push_back(h.vel[3 * i + 1]); vz.push_back(h.vel[3 * i + 2]);
                bx.push_back(b[3 * i]); by.push_back(b[3 * i + 1]); bz.push_back(b[3 * i + 2]);
            }
        }
        upload(B.X, bx, st); upload(B.Y, by, st); upload(B.Z, bz, st);
        upload(B.VX, vx, st); upload(B.VY, vy, st); upload(B.VZ, vz, st);
        request_rebuild();
        HIPCHK(hipStreamSynchronize(st));
        // first build; re-lay out with larger list capacities until the lists fit
        launch_nlist<P>(B, true);
        // positions moved on from the build positions: repack before the next forces
        if (carried) { upload(B.X, x, st); upload(B.Y, y, st); upload(B.Z, z, st); }
        pack_needed_ = carried;
        HIPCHK(hipStreamSynchronize(st));
        std::vector<SysCount> c(ns);
        copy_sync(c.data(), B.d_cnt, sizeof(SysCount) * ns, hipMemcpyDeviceToHost);
        bool grow = false;
        failed_sys_.clear();
        for (int s = 0; s < ns; ++s)
            if (c[s].overflow & 2) failed_sys_.push_back(s);
        if (!failed_sys_.empty()) {
            grow_depth_ = 0;
            throw HipError(NX_ERR_CAPACITY, "neighbour-list column exceeds 2048 atoms in system " +
                                                std::to_string(failed_sys_[0]));
        }
        std::vector<int> growing;
        for (int s = 0; s < ns; ++s) {
            if (c[s].overflow) { hs_[s]->list_scale *= 2.0; grow = true; growing.push_back(s); }
            else if (c[s].list_count > 0.8 * B.sys[s].list_cap || c[s].emask_count > 0.8 * B.sys[s].emask_cap) {
                hs_[s]->list_scale *= 1.6;
                grow = true;
                growing.push_back(s);
            }
        }
        counts_ = c;
        if (grow) {
            if (++grow_depth_ > 6) {
                // the systems that still do not fit are named as failed
                grow_depth_ = 0;
                failed_sys_ = growing;
                throw HipError(NX_ERR_CAPACITY, "pair list does not fit in system " + std::to_string(growing[0]));
            }
            std::vector<HostSystem*> keep = hs_;
            setup(keep);
        }
        grow_depth_ = 0;
    }

    // Host<->device copies go through the engine stream (it does not order
    // with the null stream), then wait.
    void copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
        HIPCHK(hipMemcpyAsync(dst, src, bytes, kind, B.stream));
        HIPCHK(hipStreamSynchronize(B.stream));
    }

    void request_rebuild() {
        std::vector<SysCount> c(B.nsys);
        copy_sync(c.data(), B.d_cnt, sizeof(SysCount) * B.nsys, hipMemcpyDeviceToHost);
        for (auto& v : c) { v.rebuild = 1; v.overflow = 0; }
        force_full_ = true;
        copy_sync(B.d_cnt, c.data(), sizeof(SysCount) * B.nsys, hipMemcpyHostToDevice);
    }

    // ---------------------------------------------------------------- steps

    // a..b spans n steps (n = 1 for a single kernel family)
    struct Ev { const char* name; hipEvent_t a, b; int n = 1; };
    hipEvent_t ev() {
        if (ev_used_ < evpool_.size()) return evpool_[ev_used_++];
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        evpool_.push_back(e);
        ev_used_++;
        return e;
    }
    template <class F>
    void timed(const char* name, F&& f) {
        if (!B.timing) { f(); return; }
        // inside a capture the records become event nodes of the graph
        hipEvent_t a = ev(), b = ev();
        HIPCHK(hipEventRecord(a, B.stream));
        f();
        HIPCHK(hipEventRecord(b, B.stream));
        evs_.push_back({name, a, b});
    }

    // Fold the elapsed times of recorded events into times_ and recycle the events.
    void harvest() {
        for (auto& e : evs_) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, e.a, e.b));
            auto& t = times_[e.name];
            t.first += ms;
            t.second += e.n;
        }
        evs_.clear();
        ev_used_ = 0;
    }

    void pme(bool energy) {
        timed("pme_spread", [&] { launch_pme_spread<P>(B); });
        timed("fft", [&] {
            launch_fft_forward<P>(B);
            launch_fft_convolve<P>(B, energy);
            launch_fft_inverse<P>(B);
        });
        timed("pme_gather", [&] { launch_pme_gather<P>(B); });
    }

    // Launches f on the side stream: the kernels it launches read B.stream.
    template <class F>
    void on_side(F&& f) {
        struct Restore {
            hipStream_t& ref;
            hipStream_t saved;
            ~Restore() { ref = saved; }
        } restore{B.stream, B.stream};
        B.stream = side_;
        f();
    }

    // Parts of one force evaluation: list and slot positions; bonded and
    // exclusion terms; the pair kernel; PME.
    enum { PART_LIST = 1, PART_TERMS = 2, PART_PAIR = 4, PART_PME = 8, PART_ALL = 15 };

    // Whether the terms and PME run on the side stream beside the pair kernel
    // (not under per-kernel timing, whose times are those of kernels alone).
    bool use_side(int mask) const {
        return overlap_ && (mask & NX_FORCE_NONBONDED) && B.pme_any && B.timing != NX_TIMING_KERNELS;
    }

    // The bonded and exclusion terms and PME (the work beside the pair kernel).
    // (Starting the terms earlier, beside the list work, measured slower: 0.263
    // against 0.240 ms per 100k-atom mixed step.)
    void side_work(int mask, bool energy) {
        if (mask & 7) timed("bonded", [&] { launch_bonded<P>(B, mask, energy); });
        timed("atom_forces", [&] { launch_atom_forces<P>(B, mask, energy); });
        if (mask & NX_FORCE_NONBONDED) pme(energy);
    }

    // The parts in `parts`.  With every part and the side stream in use, the
    // terms and PME go to the side stream once the list work has laid out the
    // slots (before its last kernel, the chunk partition, which PME does not
    // read) and run beside the pair kernel (the pair kernel and PME add to the
    // slot forces with atomics, the terms write the atom forces); the main
    // stream then waits for the side stream.
    void forces_part(int mask, bool energy, int parts) {
        const bool nb = (mask & NX_FORCE_NONBONDED) != 0;
        if (parts == PART_ALL && use_side(mask)) {
            // the side stream forks before the list's last kernel, which then does the
            // pair kernel's set-up too: one launch fewer on the main stream (0.240 ->
            // 0.237 ms per 100k-atom mixed step), and the side kernels still reach the
            // CUs before the pair kernel's persistent grid
            timed("nlist", [&] { launch_nlist<P>(B, !nb, 1); });
            settle();
            fork_side();
            on_side([&] { side_work(mask, energy); });
            timed("nlist", [&] { launch_nlist<P>(B, !nb, 2); });
            timed("pair", [&] { launch_pair_kernel<P>(B, energy); });
            join_side();
            return;
        }
        if (parts & PART_LIST) {
            // the pair kernel's first launch clears the rebuild flags
            timed("nlist", [&] { launch_nlist<P>(B, !nb); });
            settle();
        }
        if (parts & PART_TERMS) {
            if (mask & 7) timed("bonded", [&] { launch_bonded<P>(B, mask, energy); });
            timed("atom_forces", [&] { launch_atom_forces<P>(B, mask, energy); });
        }
        if (nb && (parts & PART_PAIR)) {
            launch_pair_begin<P>(B);
            timed("pair", [&] { launch_pair_kernel<P>(B, energy); });
        }
        if (nb && (parts & PART_PME)) pme(energy);
    }
    void forces(int mask, bool energy) { forces_part(mask, energy, PART_ALL); }

    // Slot positions and slot forces are kept current by the list build and the
    // update; after positions were replaced or forces computed without an update,
    // one packing pass restores them (outside any graph).
    void settle() {
        if (!pack_needed_ || capturing_) return;
        timed("pack", [&] { launch_pack<P>(B); });
        pack_needed_ = false;
    }

    // The side stream waits for the work queued so far on the main stream; the
    // main stream waits for the side stream.
    void fork_side() {
        HIPCHK(hipEventRecord(fork_, B.stream));
        HIPCHK(hipStreamWaitEvent(side_, fork_, 0));
    }
    void join_side() {
        HIPCHK(hipEventRecord(join_, side_));
        HIPCHK(hipStreamWaitEvent(B.stream, join_, 0));
    }

    void zero_energy() { HIPCHK(hipMemsetAsync(B.energy, 0, sizeof(long long) * B.nsys * E_NTERMS, B.stream)); }

    void begin_call() {
        evs_.clear();
        ev_used_ = 0;
        times_.clear();
        steps_ = 0;
        HIPCHK(hipGetLastError());
    }

    void end_call() {
        HIPCHK(hipStreamSynchronize(B.stream));
        HIPCHK(hipGetLastError());
        harvest();
        check_counters();
    }

    // Every system whose list build overflowed during the call is named (the
    // device froze each at its last completed step); the error names the first.
    void check_counters() {
        std::vector<SysCount> c(B.nsys);
        copy_sync(c.data(), B.d_cnt, sizeof(SysCount) * B.nsys, hipMemcpyDeviceToHost);
        counts_ = c;
        failed_sys_.clear();
        for (int s = 0; s < B.nsys; ++s)
            if (c[s].overflow & ~kStaleList) failed_sys_.push_back(s);
        if (!failed_sys_.empty()) {
            const int s = failed_sys_[0];
            throw HipError(NX_ERR_CAPACITY, "neighbour-list capacity exceeded in system " + std::to_string(s) +
                                                " (code " + std::to_string(c[s].overflow) + ", " +
                                                std::to_string(failed_sys_.size()) + " system(s) failed)");
        }
    }

    void step_plain(double dt, int mask) {
        forces(mask, false);
        timed("integrate", [&] { launch_integrate<P>(B, dt); });
        ++steps_;
    }

    void run(int nsteps, double dt, int mask) override {
        begin_call();
        settle();
        if (B.timing == NX_TIMING_GRAPH_PAIR && nsteps > 0 && graphs_enabled_) {
            // Every step replays as graphs around the pair kernel -- [list,
            // positions], the side graph (terms, PME; on the side stream beside the
            // pair kernel), the pair kernel, [update] -- with events on the main
            // stream around the pair kernel and one after every step.
            if (!seg_exec_[0] || graph_dt_seg_ != dt || graph_mask_seg_ != mask) build_segments(dt, mask);
            const bool side = use_side(mask);
            hipEvent_t e0 = ev();
            HIPCHK(hipEventRecord(e0, B.stream));
            for (int i = 0; i < nsteps; ++i) {
                hipEvent_t a = ev(), b = ev(), c = ev(), la = ev(), lb = ev(), ua = ev();
                HIPCHK(hipEventRecord(la, B.stream));
                HIPCHK(hipGraphLaunch(seg_exec_[SEG_FIRST], B.stream));
                HIPCHK(hipEventRecord(lb, B.stream));
                evs_.push_back({"list", la, lb});
                if (side) {
                    fork_side();
                    hipEvent_t sa = ev(), sb = ev();
                    HIPCHK(hipEventRecord(sa, side_));
                    HIPCHK(hipGraphLaunch(seg_exec_[SEG_SIDE], side_));
                    HIPCHK(hipEventRecord(sb, side_));
                    evs_.push_back({"side", sa, sb});
                } else {
                    HIPCHK(hipGraphLaunch(seg_exec_[SEG_SIDE], B.stream));
                }
                if (seg_exec_[SEG_BEGIN]) HIPCHK(hipGraphLaunch(seg_exec_[SEG_BEGIN], B.stream));
                HIPCHK(hipEventRecord(a, B.stream));
                if (seg_exec_[SEG_PAIR]) HIPCHK(hipGraphLaunch(seg_exec_[SEG_PAIR], B.stream));
                HIPCHK(hipEventRecord(b, B.stream));
                if (side) join_side();
                HIPCHK(hipEventRecord(ua, B.stream));
                HIPCHK(hipGraphLaunch(seg_exec_[SEG_LAST], B.stream));
                HIPCHK(hipEventRecord(c, B.stream));
                evs_.push_back({"update", ua, c});
                evs_.push_back({"pair", a, b});
                evs_.push_back({"step", e0, c});
                e0 = c;
            }
            steps_ += nsteps;
        } else if (B.timing) {
            // per-kernel HIP events need plain launches (events recorded in a graph
            // cannot be queried for elapsed time); the host enqueues ahead of the GPU
            for (int i = 0; i < nsteps; ++i) step_plain(dt, mask);
        } else if (nsteps >= 4 && graphs_enabled_ && lite_ok(mask)) {
            if (!graph_exec_ || graph_dt_ != dt || graph_mask_ != mask) build_graph(dt, mask);
            run_lite(nsteps);
        } else if (nsteps >= 4 && graphs_enabled_) {
            if (!graph_exec_ || graph_dt_ != dt || graph_mask_ != mask) build_graph(dt, mask);
            for (int i = 0; i < nsteps; ++i) HIPCHK(hipGraphLaunch(graph_exec_, B.stream));
            steps_ += nsteps;
        } else {
            for (int i = 0; i < nsteps; ++i) step_plain(dt, mask);
        }
        end_call();
    }

    // nsteps steps (graph replays where run() would use them) with an event on
    // the main stream before the first step and after every step; ms[i] is the
    // device time from the end of step i - 1 (the first: from the start) to the
    // end of step i.
    void run_step_times(int nsteps, double dt, int mask, std::vector<double>& ms) override {
        ms.clear();
        if (nsteps <= 0) return;
        begin_call();
        settle();
        const bool graph = nsteps >= 4 && graphs_enabled_;
        if (graph && (!graph_exec_ || graph_dt_ != dt || graph_mask_ != mask)) build_graph(dt, mask);
        if ((int)step_ev_.size() < nsteps + 1) {
            for (int i = (int)step_ev_.size(); i < nsteps + 1; ++i) {
                hipEvent_t e;
                HIPCHK(hipEventCreate(&e));
                step_ev_.push_back(e);
            }
        }
        HIPCHK(hipEventRecord(step_ev_[0], B.stream));
        for (int i = 0; i < nsteps; ++i) {
            if (graph) {
                HIPCHK(hipGraphLaunch(graph_exec_, B.stream));
                ++steps_;
            } else {
                step_plain(dt, mask);
            }
            HIPCHK(hipEventRecord(step_ev_[i + 1], B.stream));
        }
        end_call();
        ms.resize(nsteps);
        for (int i = 0; i < nsteps; ++i) {
            float t = 0.f;
            HIPCHK(hipEventElapsedTime(&t, step_ev_[i], step_ev_[i + 1]));
            ms[i] = t;
        }
    }

    void build_graph(double dt, int mask) {
        destroy_graphs();
        capturing_ = true;
        HIPCHK(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
        forces(mask, false);
        launch_integrate<P>(B, dt);
        hipError_t e = hipStreamEndCapture(B.stream, &graph_);
        if (e == hipSuccess && lite_ok(mask)) {
            // the same step without the list kernels before the fork
            e = hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                forces_lite(mask);
                launch_integrate<P>(B, dt);
                e = hipStreamEndCapture(B.stream, &graph_lite_);
            }
        }
        capturing_ = false;
        HIPCHK(e);
        HIPCHK(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
        if (graph_lite_) HIPCHK(hipGraphInstantiate(&graph_lite_exec_, graph_lite_, nullptr, nullptr, 0));
        graph_dt_ = dt;
        graph_mask_ = mask;
    }
    void destroy_graphs() {
        if (graph_exec_) { hipGraphExecDestroy(graph_exec_); graph_exec_ = nullptr; }
        if (graph_) { hipGraphDestroy(graph_); graph_ = nullptr; }
        if (graph_lite_exec_) { hipGraphExecDestroy(graph_lite_exec_); graph_lite_exec_ = nullptr; }
        if (graph_lite_) { hipGraphDestroy(graph_lite_); graph_lite_ = nullptr; }
    }

    // ---------------------------------------------- steps without the list kernels
    //
    // On most steps no list is built, and the five list kernels before the fork
    // only read the rebuild flag and exit; their launches still cost the step
    // about 15 us on the 100k-atom system.  A one-system batch therefore has a
    // second step graph without them (nl_waves still runs, after the fork).
    // The host picks one of the two before each step, from the report nl_waves
    // writes to host memory early in every step (StepInfo): the largest
    // displacement from the build positions after the update before it, and
    // its growth per step.  The
    // device keeps the result exact whatever the pick: a step without the list
    // kernels on a system whose rebuild flag is up marks it stale (nl_waves),
    // the update freezes it, and the host runs that step again with the list
    // kernels.  Every build therefore happens at the step its flag names, as
    // with the list kernels in every step, and the results are the same bits.
    bool lite_ok(int mask) const {
        return lite_mode_ && B.step_info && B.nsys == 1 && use_side(mask) && B.timing == NX_TIMING_OFF;
    }
    void forces_lite(int mask) {
        const bool nb = (mask & NX_FORCE_NONBONDED) != 0;
        fork_side();
        on_side([&] { side_work(mask, false); });
        launch_nlist<P>(B, !nb, 3);
        launch_pair_kernel<P>(B, false);
        join_side();
    }
    // the growth of the largest displacement per step, from consecutive reports
    void observe(const volatile StepInfo* in) {
        const unsigned seq = in->seq;
        const float d = in->dmax;
        if (seq == last_seq_ + 1 && d >= last_dmax_ && last_seq_ > 0) {
            grow_[grow_n_ % kGrowWindow] = d - last_dmax_;
            ++grow_n_;
        }
        if (seq != last_seq_) { last_seq_ = seq; last_dmax_ = d; }
    }
    void run_lite(int nsteps) {
        const volatile StepInfo* in = info_host_;
        const double half = std::sqrt(B.sys[0].skin_half2);
        constexpr int R = 4;        // steps whose graph choice is remembered
        std::vector<SysCount> c(1);
        HIPCHK(hipStreamSynchronize(B.stream));
        copy_sync(c.data(), B.d_cnt, sizeof(SysCount), hipMemcpyDeviceToHost);
        const int done0 = c[0].steps_done;
        long long remaining = nsteps, launched_total = 0;
        // Rounds: launch the remaining steps, choosing each graph from the report
        // of the step before it; then read the exact count of steps done (a step
        // without the list kernels that found a build due froze and is run again
        // in the next round).
        while (remaining > 0 && !(c[0].overflow & ~kStaleList)) {
            const unsigned seq0 = in->seq;       // every earlier step has reported
            bool full_of[R] = {};
            for (long long j = 0; j < remaining; ++j) {
                bool full;
                if (j == 0) {
                    full = force_full_ || c[0].rebuild || (c[0].overflow & kStaleList);     // exact
                } else {
                    // Step j is chosen from the report of step j - 1, which its nl_waves
                    // writes early in that step: graph j is then queued while step j - 1
                    // runs.  (Waiting for the report of step j - 2 instead forecast two
                    // steps ahead and ran fewer steps without the list kernels: 0.2253
                    // against 0.2227 ms per 100k-atom mixed step.)
                    if ((long long)(unsigned)(in->seq - seq0) < j) {
                        // host memory only (a stream query in the loop delayed the GPU's
                        // work by several us a step); after 0.2 s without the report the
                        // stream is asked, and a drained stream without it is an error
                        const auto t0 = std::chrono::steady_clock::now();
                        for (unsigned spin = 1; (long long)(unsigned)(in->seq - seq0) < j; ++spin) {
                            if ((spin & 4095u) != 0) continue;
                            if (std::chrono::steady_clock::now() - t0 < std::chrono::milliseconds(200)) continue;
                            const hipError_t q = hipStreamQuery(B.stream);
                            if (q == hipSuccess && (long long)(unsigned)(in->seq - seq0) < j)
                                throw HipError(NX_ERR_STATE, "a finished step left no report in host memory");
                            if (q != hipSuccess && q != hipErrorNotReady) HIPCHK(q);
                        }
                    }
                    observe(in);
                    const bool due = (in->flags & 1) != 0, prev_full = full_of[(j - 1) % R];
                    if (due && !prev_full) {
                        full = true;             // step j - 1 freezes; step j builds
                    } else if (lite_mode_ == 2) {
                        full = false;
                    } else {
                        // step j reads the positions after step j - 1's update: the
                        // displacement of the report (after the update before step j - 1;
                        // zero if step j - 1 builds) plus one step's growth, taken as
                        // twice the largest of the recent steps, plus a margin
                        float g = 0.0f;
                        const int n = std::min(grow_n_, kGrowWindow);
                        for (int i = 0; i < n; ++i) g = std::max(g, grow_[i]);
                        const double base = due ? 0.0 : (double)in->dmax;
                        full = n < 2 || base + 2.0 * g + 0.02 * half >= half;
                    }
                }
                HIPCHK(hipGraphLaunch(full ? graph_exec_ : graph_lite_exec_, B.stream));
                full_of[j % R] = full;
                if (full) force_full_ = false;
                if (full) ++lite_full_; else ++lite_lite_;
                ++launched_total;
            }
            HIPCHK(hipStreamSynchronize(B.stream));
            copy_sync(c.data(), B.d_cnt, sizeof(SysCount), hipMemcpyDeviceToHost);
            const long long left = nsteps - (c[0].steps_done - done0);
            if (left >= remaining && !(c[0].overflow & ~kStaleList))
                throw HipError(NX_ERR_STATE, "a round of steps made no progress");
            remaining = left;
        }
        const long long executed = c[0].steps_done - done0;
        lite_repeats_ += launched_total - executed;
        steps_ += executed;
    }

    void build_segments(double dt, int mask) {
        destroy_segments();
        const int saved = B.timing;
        B.timing = 0;
        capturing_ = true;
        for (int g = 0; g < SEG_COUNT; ++g) {
            if ((g == SEG_PAIR || g == SEG_BEGIN) && !(mask & NX_FORCE_NONBONDED)) continue;     // no pair graphs
            HIPCHK(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
            if (g == SEG_LAST) launch_integrate<P>(B, dt);
            if (g == SEG_FIRST) forces_part(mask, false, PART_LIST);
            if (g == SEG_SIDE) side_work(mask, false);
            if (g == SEG_BEGIN) launch_pair_begin<P>(B);
            if (g == SEG_PAIR) launch_pair_kernel<P>(B, false);
            hipError_t e = hipStreamEndCapture(B.stream, &seg_[g]);
            if (e != hipSuccess) { capturing_ = false; B.timing = saved; HIPCHK(e); }
            HIPCHK(hipGraphInstantiate(&seg_exec_[g], seg_[g], nullptr, nullptr, 0));
        }
        capturing_ = false;
        B.timing = saved;
        graph_dt_seg_ = dt;
        graph_mask_seg_ = mask;
    }
    void destroy_segments() {
        for (int g = 0; g < SEG_COUNT; ++g) {
            if (seg_exec_[g]) { hipGraphExecDestroy(seg_exec_[g]); seg_exec_[g] = nullptr; }
            if (seg_[g]) { hipGraphDestroy(seg_[g]); seg_[g] = nullptr; }
        }
    }

    void energies_out(int s, int mask, nx_energies* e) {
        std::vector<long long> h(E_NTERMS);
        copy_sync(h.data(), B.energy + (size_t)s * E_NTERMS, sizeof(long long) * E_NTERMS, hipMemcpyDeviceToHost);
        energies_from(h.data(), s, mask, e);
    }

    // system s's energies from its E_NTERMS fixed-point sums h
    void energies_from(const long long* h, int s, int mask, nx_energies* e) {
        auto v = [&](int t) { return (double)h[t] / kEnergyScale; };
        std::memset(e, 0, sizeof(*e));
        e->bond = (mask & 1) ? v(E_BOND) : 0.0;
        e->angle = (mask & 2) ? v(E_ANGLE) : 0.0;
        e->torsion = (mask & 4) ? v(E_TORSION) : 0.0;
        e->kinetic = v(E_KIN);
        const double bonded = e->bond + e->angle + e->torsion;
        e->potential = bonded;
        e->reference_pe = bonded;
        if (mask & NX_FORCE_NONBONDED) {
            e->coulomb_direct = v(E_COUL);
            e->lj = v(E_LJ);
            e->recip = v(E_RECIP);
            e->self = self_energy_[s];
            e->exclusion = v(E_EXCL);
            e->reference_pe = e->recip + e->self + (e->coulomb_direct + e->lj) - e->exclusion;
            e->potential = bonded + e->recip + e->self + e->coulomb_direct + e->lj - e->exclusion;
        }
    }

    // Forces + energies at the current positions; optionally kinetic energy and integrate.
    void compute(int mask, bool kinetic, double dt, bool integrate, std::vector<double>* F,
                 std::vector<double>* pos, std::vector<double>* vel, std::vector<nx_energies>* E) override {
        begin_call();
        zero_energy();
        forces(mask, true);
        if (kinetic) launch_kinetic<P>(B, dt);
        HIPCHK(hipStreamSynchronize(B.stream));
        if (F) gather_forces(*F);
        if (pos || vel) get_state(pos, vel);
        if (E) {
            E->resize(B.nsys);
            for (int s = 0; s < B.nsys; ++s) energies_out(s, mask, &(*E)[s]);
        }
        if (integrate) {
            timed("integrate", [&] { launch_integrate<P>(B, dt); });
            ++steps_;
        } else {
            pack_needed_ = true;          // the slot forces still hold this evaluation
        }
        end_call();
    }

    // ------------------------------------------------------ async reports
    //
    // One report step without a host round trip: forces and energies at x(t),
    // the kinetic energy with the half-step shift, a snapshot of x(t),
    // v(t - dt/2) and F(t) into a device staging block of the slot, then the
    // leapfrog update, all on the engine stream.  A copy stream moves the
    // snapshot, the energy sums and the list counters to the caller's (pinned)
    // buffers after an event, so the update and the next steps do not wait for
    // the copies.  report_wait(ticket) waits for the slot's copies and returns
    // its energies.  A slot's staging is reused only after its ticket was
    // waited for, so no copy can read a block that a later report writes.
    long long report_async(int mask, double dt, double* pos, double* vel, double* F) override {
        int k = -1;
        for (int i = 0; i < kReportSlots; ++i)
            if (rs_[i].ticket < 0) { k = i; break; }
        if (k < 0)
            throw HipError(NX_ERR_STATE, "too many reports in flight (" + std::to_string(kReportSlots) +
                                             "): wait for one with nx_report_wait first");
        ReportSlot& r = rs_[k];
        const size_t n = B.n_atoms;
        const size_t ebytes = sizeof(long long) * B.nsys * E_NTERMS, cbytes = sizeof(SysCount) * B.nsys;
        if (!copy_) HIPCHK(hipStreamCreateWithFlags(&copy_, hipStreamNonBlocking));
        if (!packed_) HIPCHK(hipEventCreateWithFlags(&packed_, hipEventDisableTiming));
        if (!r.d_stage) {
            HIPCHK(hipMalloc(&r.d_stage, 9 * n * sizeof(double)));
            HIPCHK(hipMalloc(&r.d_energy, ebytes));
            HIPCHK(hipMalloc(&r.d_cnt, cbytes));
            HIPCHK(hipHostMalloc(&r.h_energy, ebytes, hipHostMallocDefault));
            HIPCHK(hipHostMalloc(&r.h_cnt, cbytes, hipHostMallocDefault));
            HIPCHK(hipEventCreateWithFlags(&r.done, hipEventDisableTiming));
        }
        evs_.clear();
        ev_used_ = 0;
        HIPCHK(hipGetLastError());
        settle();
        zero_energy();
        forces(mask, true);
        launch_kinetic<P>(B, dt);
        launch_report_snapshot<P>(B, r.d_stage);
        HIPCHK(hipMemcpyAsync(r.d_energy, B.energy, ebytes, hipMemcpyDeviceToDevice, B.stream));
        HIPCHK(hipMemcpyAsync(r.d_cnt, B.d_cnt, cbytes, hipMemcpyDeviceToDevice, B.stream));
        HIPCHK(hipEventRecord(packed_, B.stream));
        HIPCHK(hipStreamWaitEvent(copy_, packed_, 0));
        double* dst[3] = {pos, vel, F};
        for (int b = 0; b < 3; ++b)
            if (dst[b])
                HIPCHK(hipMemcpyAsync(dst[b], r.d_stage + 3 * n * b, 3 * n * sizeof(double), hipMemcpyDeviceToHost,
                                      copy_));
        HIPCHK(hipMemcpyAsync(r.h_energy, r.d_energy, ebytes, hipMemcpyDeviceToHost, copy_));
        HIPCHK(hipMemcpyAsync(r.h_cnt, r.d_cnt, cbytes, hipMemcpyDeviceToHost, copy_));
        HIPCHK(hipEventRecord(r.done, copy_));
        launch_integrate<P>(B, dt);
        ++steps_;
        pack_needed_ = false;
        r.mask = mask;
        r.ticket = next_ticket_++;
        return r.ticket;
    }

    void report_wait(long long ticket, std::vector<nx_energies>* E) override {
        int k = -1;
        for (int i = 0; i < kReportSlots; ++i)
            if (rs_[i].ticket == ticket && ticket >= 0) k = i;
        if (k < 0) throw HipError(NX_ERR_ARG, "unknown or already collected report ticket " + std::to_string(ticket));
        ReportSlot& r = rs_[k];
        const hipError_t e = hipEventSynchronize(r.done);
        r.ticket = -1;
        HIPCHK(e);
        HIPCHK(hipGetLastError());
        counts_.assign(r.h_cnt, r.h_cnt + B.nsys);
        failed_sys_.clear();
        for (int s = 0; s < B.nsys; ++s)
            if (r.h_cnt[s].overflow) failed_sys_.push_back(s);
        if (!failed_sys_.empty())
            throw HipError(NX_ERR_CAPACITY, "neighbour-list capacity exceeded in system " +
                                                std::to_string(failed_sys_[0]) + " (code " +
                                                std::to_string(r.h_cnt[failed_sys_[0]].overflow) + ")");
        if (E) {
            E->resize(B.nsys);
            for (int s = 0; s < B.nsys; ++s) energies_from(r.h_energy + (size_t)s * E_NTERMS, s, r.mask, &(*E)[s]);
        }
    }

    // waits for every report in flight (their results are dropped) and frees the slots
    void drain_reports() {
        for (auto& r : rs_) {
            if (r.done && r.ticket >= 0) hipEventSynchronize(r.done);
            r.ticket = -1;
        }
    }
    void free_reports() {
        drain_reports();
        for (auto& r : rs_) {
            if (r.d_stage) hipFree(r.d_stage);
            if (r.d_energy) hipFree(r.d_energy);
            if (r.d_cnt) hipFree(r.d_cnt);
            if (r.h_energy) hipHostFree(r.h_energy);
            if (r.h_cnt) hipHostFree(r.h_cnt);
            if (r.done) hipEventDestroy(r.done);
            r = ReportSlot{};
        }
    }

    void gather_forces(std::vector<double>& F) {
        double* d;
        HIPCHK(hipMallocAsync(reinterpret_cast<void**>(&d), sizeof(double) * 3 * (size_t)B.n_atoms, B.stream));
        launch_collect_forces<P>(B, d, d + B.n_atoms, d + 2 * (size_t)B.n_atoms);
        std::vector<double> h(3 * (size_t)B.n_atoms);
        HIPCHK(hipMemcpyAsync(h.data(), d, h.size() * sizeof(double), hipMemcpyDeviceToHost, B.stream));
        HIPCHK(hipFreeAsync(d, B.stream));
        HIPCHK(hipStreamSynchronize(B.stream));
        F.resize(3 * (size_t)B.n_atoms);
        for (int a = 0; a < B.n_atoms; ++a)
            for (int k = 0; k < 3; ++k) F[3 * (size_t)a + k] = h[(size_t)k * B.n_atoms + a];
    }

    void get_state(std::vector<double>* pos, std::vector<double>* vel) override {
        const size_t n = B.n_atoms;
        std::vector<double> h(3 * n);
        if (pos) {
            HIPCHK(hipMemcpyAsync(h.data(), B.X, n * 8, hipMemcpyDeviceToHost, B.stream));
            HIPCHK(hipMemcpyAsync(h.data() + n, B.Y, n * 8, hipMemcpyDeviceToHost, B.stream));
            HIPCHK(hipMemcpyAsync(h.data() + 2 * n, B.Z, n * 8, hipMemcpyDeviceToHost, B.stream));
            HIPCHK(hipStreamSynchronize(B.stream));
            pos->resize(3 * n);
            for (size_t a = 0; a < n; ++a)
                for (int k = 0; k < 3; ++k) (*pos)[3 * a + k] = h[k * n + a];
        }
        if (vel) {
            HIPCHK(hipMemcpyAsync(h.data(), B.VX, n * 8, hipMemcpyDeviceToHost, B.stream));
            HIPCHK(hipMemcpyAsync(h.data() + n, B.VY, n * 8, hipMemcpyDeviceToHost, B.stream));
            HIPCHK(hipMemcpyAsync(h.data() + 2 * n, B.VZ, n * 8, hipMemcpyDeviceToHost, B.stream));
            HIPCHK(hipStreamSynchronize(B.stream));
            vel->resize(3 * n);
            for (size_t a = 0; a < n; ++a)
                for (int k = 0; k < 3; ++k) (*vel)[3 * a + k] = h[k * n + a];
        }
    }

    void set_state(const double* pos, const double* vel) override {
        const size_t n = B.n_atoms;
        std::vector<double> h(3 * n);
        for (int k = 0; k < 3; ++k)
            for (size_t a = 0; a < n; ++a) h[k * n + a] = pos[3 * a + k];
        copy_sync(B.X, h.data(), n * 8, hipMemcpyHostToDevice);
        copy_sync(B.Y, h.data() + n, n * 8, hipMemcpyHostToDevice);
        copy_sync(B.Z, h.data() + 2 * n, n * 8, hipMemcpyHostToDevice);
        for (int k = 0; k < 3; ++k)
            for (size_t a = 0; a < n; ++a) h[k * n + a] = vel ? vel[3 * a + k] : 0.0;
        copy_sync(B.VX, h.data(), n * 8, hipMemcpyHostToDevice);
        copy_sync(B.VY, h.data() + n, n * 8, hipMemcpyHostToDevice);
        copy_sync(B.VZ, h.data() + 2 * n, n * 8, hipMemcpyHostToDevice);
        request_rebuild();
    }

    long long export_pairs(std::vector<int>* out) override {
        begin_call();
        // make the list current for these positions
        launch_nlist<P>(B, true);
        HIPCHK(hipStreamSynchronize(B.stream));
        check_counters();
        long long cap = out ? (1 << 20) : 0;
        while (true) {
            int2* d = nullptr;
            HIPCHK(hipMalloc(&d, sizeof(int2) * std::max<long long>(cap, 1)));
            std::unique_ptr<int2, hipError_t (*)(void*)> hold(d, hipFree);   // freed on every path out
            HIPCHK(hipMemsetAsync(d_count_, 0, sizeof(long long), B.stream));
            launch_export_pairs<P>(B, d, cap, d_count_);
            HIPCHK(hipGetLastError());
            long long n = 0;
            HIPCHK(hipMemcpyAsync(&n, d_count_, sizeof(long long), hipMemcpyDeviceToHost, B.stream));
            HIPCHK(hipStreamSynchronize(B.stream));
            if (!out) return n;
            if (n <= cap) {
                std::vector<int2> h(n);
                if (n) copy_sync(h.data(), d, sizeof(int2) * n, hipMemcpyDeviceToHost);
                out->resize(2 * n);
                for (long long i = 0; i < n; ++i) { (*out)[2 * i] = h[i].x; (*out)[2 * i + 1] = h[i].y; }
                return n;
            }
            cap = n;
        }
    }

    void stats(nx_stats* st) override {
        std::memset(st, 0, sizeof(*st));
        std::vector<SysCount> c(B.nsys);
        copy_sync(c.data(), B.d_cnt, sizeof(SysCount) * B.nsys, hipMemcpyDeviceToHost);
        long long tiles = 0;
        std::vector<int4> items(B.n_super_cap);
        copy_sync(items.data(), B.super_item, sizeof(int4) * B.n_super_cap, hipMemcpyDeviceToHost);
        std::vector<long long> off(B.nsys + 1, 0);
        for (int s = 0; s < B.nsys; ++s) {
            st->n_slots += c[s].n_slots;
            st->n_clusters += c[s].n_clusters;
            st->n_super += c[s].n_super;
            st->list_entries += c[s].list_count;
            st->rebuilds += c[s].rebuilds;
            for (int i = 0; i < c[s].n_super; ++i) (void)i;
        }
        // tiles = sum over entries of popcount(imask): read the list regions
        for (int s = 0; s < B.nsys; ++s) {
            const SysDev& d = B.sys[s];
            std::vector<uint2> l(c[s].list_count);
            if (!l.empty())
                copy_sync(l.data(), B.list + d.list_off, sizeof(uint2) * l.size(), hipMemcpyDeviceToHost);
            for (auto& v : l) tiles += __builtin_popcount(v.x >> 24);
        }
        st->tiles = tiles;
        st->steps = steps_;
        for (int k = 0; k < 3; ++k) st->grid[k] = B.sys.empty() ? 0 : B.sys[0].grid[k];
        st->alpha = B.sys.empty() ? 0 : B.sys[0].alpha;
        st->steps_without_list = lite_lite_;
        st->steps_repeated = lite_repeats_;
    }

    bool kernel_time(const std::string& name, double* ms, long long* n) override {
        auto it = times_.find(name);
        if (it == times_.end()) return false;
        *ms = it->second.second ? it->second.first / it->second.second : 0.0;
        *n = it->second.second;
        return true;
    }

    // ------------------------------------------------------------ tuning

    void set_tuning(const Tuning& t) override {
        tuning_ = t;
        tuning_set_ = t.pair_per_cu || t.spread_block || t.fft_xblocks || t.fft_lines;
    }
    Tuning tuning() const override { return tuning_; }

    // What the launch shapes depend on: precision, the device's CU count and, per
    // system, its atom count, box (to 0.05 nm), PME grid and cutoff.  Batches of
    // more than four systems are not cached.
    std::string tune_key() const {
        if (hs_.empty() || hs_.size() > 4) return std::string();
        char buf[160];
        int cu = 0;
        hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device_);
        std::string k = std::string(P::mixed ? "mixed" : "fp64") + "|cu" + std::to_string(cu);
        for (const HostSystem* h : hs_) {
            int g[3];
            for (int a = 0; a < 3; ++a) g[a] = h->nb.grid[a] > 0 ? h->nb.grid[a] : grid_rule(fabs(h->ub[a] - h->lb[a]));
            std::snprintf(buf, sizeof(buf), "|n%d,L%.0f,%.0f,%.0f,g%dx%dx%d,rc%.3f,sk%.3f", h->n,
                          20.0 * fabs(h->ub[0] - h->lb[0]), 20.0 * fabs(h->ub[1] - h->lb[1]),
                          20.0 * fabs(h->ub[2] - h->lb[2]), g[0], g[1], g[2], h->nb.cutoff, h->nb.skin);
            k += buf;
        }
        return k;
    }

    Tuning autotune(int steps, double* ms_default, double* ms_best) override {
        if (B.nsys == 0 || hs_.empty()) throw HipError(NX_ERR_STATE, "nothing laid out to tune");
        if (steps < 4) throw HipError(NX_ERR_ARG, "autotune needs at least 4 steps per trial");
        const int mask = NX_FORCE_ALL;
        const double dt = 0.0005;
        // every trial lays the batch out again from the current state, with the
        // list built from the same positions, and the state is restored after
        for (int s = 0; s < B.nsys; ++s) {
            HostSystem& h = *hs_[s];
            device_state(s, h.pos.data(), h.vel.data());
            h.build_pos.resize(h.pos.size());
            int due = 0;
            device_list_state(s, h.build_pos.data(), &due);
            h.rebuild_pending = due != 0;
        }
        const std::vector<HostSystem*> systems = hs_;
        auto trial = [&](const Tuning& t) {
            tuning_ = t;
            tuning_set_ = true;
            setup(systems);
            run(4, dt, mask);                        // graph capture and warm-up
            const auto t0 = std::chrono::steady_clock::now();
            run(steps, dt, mask);
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            return ms / steps;
        };
        Tuning best{};
        double best_ms = trial(best);
        auto consider = [&](Tuning t) {
            const double ms = trial(t);
            if (ms < 0.99 * best_ms) { best_ms = ms; best = t; }
        };
        const int occ = std::min(5, pair_blocks_occupancy<P>());
        {
            const Tuning base = best;
            for (int c = 1; c <= occ; ++c) { Tuning t = base; t.pair_per_cu = c; consider(t); }
        }
        {
            int bmax = 1;
            for (const SysDev& d : B.sys) bmax = std::max(bmax, std::min(12, std::max(d.ncx, d.ncy)));
            const Tuning base = best;
            for (int b = 1; b <= bmax; ++b) { Tuning t = base; t.spread_block = b; consider(t); }
        }
        {
            const Tuning base = best;
            if (B.fft_plane) {
                for (int x : {128, 256, 1024, 2048}) { Tuning t = base; t.fft_xblocks = x; consider(t); }
            } else {
                for (int l : {1, 2, 4}) { Tuning t = base; t.fft_lines = l; consider(t); }
            }
        }
        // the two finalists measured once more, side by side
        const double d_ms = trial(Tuning{});
        const double b_ms = trial(best);
        if (b_ms > d_ms) best = Tuning{};
        if (ms_default) *ms_default = d_ms;
        if (ms_best) *ms_best = std::min(b_ms, d_ms);
        tuning_ = best;
        tuning_set_ = true;
        setup(systems);                              // the state as it was, with the chosen shapes
        tune_cache_put(tune_key(), best);
        return best;
    }

    void set_timing(int mode) override { B.timing = mode; }
    void set_graphs(bool on) override { graphs_enabled_ = on; }
    int nsys() const override { return B.nsys; }
    const std::vector<int>& failed_systems() const override { return failed_sys_; }
    const std::vector<SysCount>& counters() const override { return counts_; }
    hipStream_t stream() const override { return B.stream; }
    void device_state(int s, double* pos, double* vel) override {
        const SysDev& d = B.sys[s];
        const size_t n = d.n_atoms, o = d.atom_off;
        std::vector<double> h(n);
        double* dst[6] = {pos, pos, pos, vel, vel, vel};
        double* src[6] = {B.X, B.Y, B.Z, B.VX, B.VY, B.VZ};
        for (int k = 0; k < 6; ++k) {
            if (!dst[k]) continue;
            HIPCHK(hipMemcpyAsync(h.data(), src[k] + o, n * 8, hipMemcpyDeviceToHost, B.stream));
            HIPCHK(hipStreamSynchronize(B.stream));
            for (size_t a = 0; a < n; ++a) dst[k][3 * a + (k % 3)] = h[a];
        }
    }
    void device_list_state(int s, double* build_pos, int* rebuild_due) override {
        const SysDev& d = B.sys[s];
        const size_t n = d.n_atoms, o = d.atom_off;
        std::vector<double> h(n);
        const double* src[3] = {B.XB, B.YB, B.ZB};
        for (int k = 0; k < 3; ++k) {
            copy_sync(h.data(), src[k] + o, n * 8, hipMemcpyDeviceToHost);
            for (size_t a = 0; a < n; ++a) build_pos[3 * a + k] = h[a];
        }
        SysCount c;
        copy_sync(&c, B.d_cnt + s, sizeof(SysCount), hipMemcpyDeviceToHost);
        *rebuild_due = c.rebuild;
    }
    size_t bytes() const override { return footprint_bytes(B); }
    static size_t footprint_bytes(const Batch<P>& b) {
        size_t A = b.n_atoms, S = b.n_slots_cap;
        return A * (14 * 8 + 3 * 4 + 3 * sizeof(facc)) + S * (3 * 4 + sizeof(real4) + sizeof(real2) + 3 * sizeof(facc)) +
               (size_t)b.n_list_cap * 8 + (size_t)b.n_emask_cap * 64 + (size_t)b.rgrid_total * (sizeof(typename P::gacc) + sizeof(real)) +
               (size_t)b.cgrid_total * (sizeof(real2) + sizeof(real));
    }

   private:
    int device_;
    Batch<P> B;
    std::vector<void*> owned_;
    std::vector<HostSystem*> hs_;
    std::vector<double> self_energy_;
    int fft_maxK_ = 0;
    long long* d_count_ = nullptr;
    std::vector<hipEvent_t> evpool_;
    std::vector<hipEvent_t> step_ev_;       // run_step_times: one event per step boundary
    Tuning tuning_;                         // launch shapes of the current layout
    bool tuning_set_ = false;               // set explicitly or by autotune (else: from the cache)
    // async reports: staging on the device, energy sums and counters in pinned memory
    static constexpr int kReportSlots = 4;
    struct ReportSlot {
        double* d_stage = nullptr;          // [3][n][3]: positions, velocities, forces
        long long* d_energy = nullptr;
        SysCount* d_cnt = nullptr;
        long long* h_energy = nullptr;
        SysCount* h_cnt = nullptr;
        hipEvent_t done = nullptr;          // the slot's copies have landed
        long long ticket = -1;              // -1: free
        int mask = 0;
    };
    ReportSlot rs_[kReportSlots];
    hipStream_t copy_ = nullptr;
    hipEvent_t packed_ = nullptr;
    long long next_ticket_ = 1;
    size_t ev_used_ = 0;
    std::vector<Ev> evs_;
    std::map<std::string, std::pair<double, long long>> times_;
    long long steps_ = 0;
    bool capturing_ = false;
    bool graphs_enabled_ = true;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    hipGraph_t graph_lite_ = nullptr;       // the step without the list kernels before the fork
    hipGraphExec_t graph_lite_exec_ = nullptr;
    int lite_mode_ = 1;                     // 0 off, 1 forecast, 2 no forecast (tests)
    StepInfo* info_host_ = nullptr;         // pinned, mapped: the update's report
    static constexpr int kGrowWindow = 16;
    float grow_[kGrowWindow] = {};
    int grow_n_ = 0;
    unsigned last_seq_ = 0;
    float last_dmax_ = 0.0f;
    bool force_full_ = true;                // the next step must carry the list kernels
    long long lite_full_ = 0, lite_lite_ = 0, lite_repeats_ = 0;   // step graphs launched, steps run again
    double graph_dt_ = 0.0;
    int graph_mask_ = -1;
    std::vector<int> failed_sys_;           // systems whose list overflowed in the last call
    int grow_depth_ = 0;
    // timing segments: [list, positions], [terms, PME], [pair], [update]
    enum { SEG_FIRST, SEG_SIDE, SEG_BEGIN, SEG_PAIR, SEG_LAST, SEG_COUNT };
    hipGraph_t seg_[SEG_COUNT] = {};
    hipGraphExec_t seg_exec_[SEG_COUNT] = {};
    hipStream_t side_ = nullptr;            // PME beside the pair kernel
    hipEvent_t fork_ = nullptr, join_ = nullptr;
    bool overlap_ = true;
    bool pair_image_ = true;
    bool pack_needed_ = false;              // slot positions/forces stale (see settle)
    double graph_dt_seg_ = 0.0;
    int graph_mask_seg_ = -1;
    std::vector<SysCount> counts_;
};

EngineBase* make_engine(int device, int precision) {
    if (precision == NX_PRECISION_MIXED) return new Engine<PrecMixed>(device);
    return new Engine<PrecFP64>(device);
}

}  // namespace nx

// Context, set-up and the C ABI of libspilady_hip.so (include/spilady_hip.h).
//
// The host side replaces the reference's global state (global.h:53-214) with
// one context per run, keeps the reference's call order (core.cpp:31-76,
// spilady.cpp:43-57) and issues every device step on the context's stream
// without host synchronisation; capacity and finiteness flags are read back
// once per sld_step call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/spilady_hip.h"
#include "sld_kernels.h"
#include "sld_model.h"
#include "sld_tables.h"

using namespace sld;

namespace {

struct LevelSet {
    std::vector<int> cells;   // cells grouped by level
    std::vector<int> off;     // level offsets (size nlev + 1)
    int* d_cells = nullptr;
    int* d_off = nullptr;
    int nlev() const { return (int)off.size() - 1; }
};

struct HostCells {
    int nx, ny, nz, C;
};

// host-side record of one replica
struct ReplicaHost {
    sld_params prm{};
    HostCells hc{};
    std::vector<int> group_size;
    std::vector<int> group_cells;  // concatenated in group order (local cell ids)
    double vol = 0.0, dx2 = 0.0, dy2 = 0.0, dz2 = 0.0;
};

}  // namespace

struct sld_ctx {
    sld_params prm{};               // parameters shared by all replicas
    int R = 1;                      // replicas
    int sel = 0;                    // replica that the state accessors act on
    std::vector<ReplicaHost> rh;    // [R]
    std::vector<Rep> reps;          // [R], mirrored in D.rep
    std::vector<int> atom_seg;      // 256-slot blocks of each replica [R + 1]
    std::vector<int> cell8_seg;     // 8-cell blocks of each replica [R + 1]
    int *d_atom_seg = nullptr, *d_cell8_seg = nullptr;
    Phys phys{};
    Tables tab{};
    Dev D{};
    hipStream_t stream = nullptr;
    // graph_mode 0: eager launches (fastest on MI355X: the per-launch cost is
    // on the device side); 1: core_dTe(dt) issues the same launches on every
    // call with the same dt, so it is captured once into a graph and
    // replayed; 2: the whole step is, one graph per cell-list buffer parity
    int graph_mode = 0;
    hipGraphExec_t dte_exec = nullptr;
    double dte_exec_dt = 0.0;
    hipGraphExec_t step_exec[2] = {nullptr, nullptr};
    int* cl_first = nullptr;        // D.cl at creation (buffer parity of a step)
    uint64_t step_npass = 0;        // noise passes of one step
    bool capturing = false;
    int device = 0;
    LevelSet spin_fwd, spin_bwd, dte_fwd, dte_bwd;  // over all replicas
    // buffers
    double *d_prow = nullptr, *d_rhotab = nullptr;
    int *d_cl2 = nullptr, *d_cn2 = nullptr;  // second cell-list buffer
    double* d_part = nullptr;                // block partials
    double* d_scal = nullptr;                // reduced scalars
    sld_obs* d_obs = nullptr;                // observables ring
    int obs_cap = 0;
    double* d_einit = nullptr;               // initial_ave_energy and err
    double total_time = 0.0;
    uint64_t pass = 0;                       // noise pass counter
    uint64_t pass_dev = 0;                   // value of the device counter D.pass_base
    uint64_t qcalls = 0;                     // core_dp_A calls so far (quantum thermostat)
    // barostats and the last check()'s stresses and pressure they act on
    sld_barostat baro{};
    double last_t_stress = 0.0, last_t_pressure = 0.0;
    sld_obs last_obs{};
    bool nb_dirty = true;
    double last_ms = 0.0;
    // per-kernel event timing of one named kernel
    std::string timed_name;
    std::vector<hipEvent_t> ev_pool;
    int ev_used = 0;
    std::vector<int> ev_launches;  // launches covered by each event pair
    bool ev_span = false;          // an event pair spans a run of launches
    int timed_launches = 0;
    double timed_ms = 0.0;
    std::string err;
    std::vector<void*> allocs;
    std::vector<size_t> alloc_bytes;  // size of each entry of allocs
    // launch shapes of the kernels whose results do not depend on them
    // (defaults: the autotuner's choice on the 2M-atom step on MI355X)
    sld_shapes shapes{64, 64, 256, 2, 256, 2048, 4};
    // parametric potential model over the replicas' frozen configurations
    struct Model {
        bool ready = false;
        double theta[sld::kNParam] = {};
        sld::ModelTab tab{};
        double *d_r = nullptr, *d_rd = nullptr, *d_h = nullptr;
        double *d_rho = nullptr, *d_fp = nullptr, *d_Q = nullptr;
        double *d_F = nullptr, *d_H = nullptr, *d_gE = nullptr, *d_gF = nullptr, *d_gH = nullptr;
        double *d_E = nullptr, *d_grad = nullptr, *d_part = nullptr;
        int* d_coff = nullptr;
        int ntot = 0;
    } model;
};

namespace {

int set_err(sld_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(ctx, call)                                                                  \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return set_err((ctx), SLD_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

template <class T>
int dalloc(sld_ctx* c, T** p, size_t n) {
    void* q = nullptr;
    HIPCHK(c, hipMalloc(&q, n * sizeof(T) + 64));
    HIPCHK(c, hipMemsetAsync(q, 0, n * sizeof(T) + 64, c->stream));
    c->allocs.push_back(q);
    c->alloc_bytes.push_back(n * sizeof(T) + 64);
    *p = (T*)q;
    return SLD_OK;
}

inline int nblk(int n, int bs) { return (n + bs - 1) / bs; }
// grid of a kernel that maps its blocks with xcd_block: a multiple of the 8 XCDs
inline int xcd_grid(int nlogical) { return (nlogical + 7) / 8 * 8; }

// Event pair around a launch of the timed kernel (see sld_time_kernel).
hipEvent_t* timing_slot(sld_ctx* c, const char* name) {
    if (c->ev_span || c->timed_name.empty()) return nullptr;
    // a template instance (k_spin_phase<4>) is timed under its kernel's name
    const size_t ln = std::strcspn(name, "<");
    if (c->timed_name.size() != ln || c->timed_name.compare(0, ln, name, ln) != 0) return nullptr;
    if (c->ev_used + 2 > (int)c->ev_pool.size()) {
        if (c->ev_pool.size() >= 8192) return nullptr;
        for (int k = 0; k < 64; ++k) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            c->ev_pool.push_back(e);
        }
    }
    hipEvent_t* ev = &c->ev_pool[c->ev_used];
    c->ev_launches.resize(c->ev_used / 2 + 1);
    c->ev_launches[c->ev_used / 2] = 1;
    c->ev_used += 2;
    return ev;
}

#define LAUNCH(ctx, kern, grid, block, ...)                                              \
    do {                                                                                 \
        if ((grid) > 0) {                                                                \
            hipEvent_t* ev_ = timing_slot((ctx), #kern);                                 \
            if (ev_) (void)hipEventRecord(ev_[0], (ctx)->stream);                        \
            hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, (ctx)->stream, __VA_ARGS__); \
            hipError_t e_ = hipGetLastError();                                           \
            if (e_ != hipSuccess)                                                        \
                return set_err((ctx), SLD_ERR_HIP, "launch %s: %s", #kern, hipGetErrorString(e_)); \
            if (ev_) (void)hipEventRecord(ev_[1], (ctx)->stream);                        \
        }                                                                                \
    } while (0)

#define TRY(x)                     \
    do {                           \
        int rc_ = (x);             \
        if (rc_ != SLD_OK) return rc_; \
    } while (0)

// ---------------------------------------------------------------------------
// host-side set-up: cells, lists, groups, levels

// map_cells.cpp:41-60
HostCells cell_counts(const double* d, double mlc) {
    double a1 = std::sin(std::atan2(d[2], d[1]));
    double a2 = std::sin(std::atan2(d[5], d[3]));
    double a3 = std::sin(std::atan2(d[5], d[4]));
    double ex = std::max(mlc / a1, mlc / a2);
    double ey = std::max(mlc, mlc / a3);
    double ez = mlc;
    HostCells h;
    h.nx = (int)(d[0] / ex);
    h.ny = (int)(d[2] / ey);
    h.nz = (int)(d[5] / ez);
    h.C = h.nx * h.ny * h.nz;
    return h;
}

void host_neigh(const HostCells& h, int c, int out[26]) {
    static const int off[26][3] = {
        {1, 0, 0},  {1, 1, 0},  {0, 1, 0},  {-1, 1, 0},  {1, 0, -1}, {1, 1, -1}, {0, 1, -1},
        {-1, 1, -1}, {1, 0, 1}, {1, 1, 1},  {0, 1, 1},   {-1, 1, 1}, {0, 0, 1},  {-1, 0, 0},
        {0, -1, 0}, {0, 0, -1}, {1, -1, 0}, {-1, -1, 0}, {-1, 0, 1}, {-1, 0, -1}, {0, -1, 1},
        {0, -1, -1}, {1, -1, 1}, {1, -1, -1}, {-1, -1, 1}, {-1, -1, -1}};
    int ix = c % h.nx, iy = (c / h.nx) % h.ny, iz = c / (h.nx * h.ny);
    for (int s = 0; s < 26; ++s) {
        int x = (ix + off[s][0] + h.nx) % h.nx, y = (iy + off[s][1] + h.ny) % h.ny,
            z = (iz + off[s][2] + h.nz) % h.nz;
        out[s] = x + y * h.nx + z * h.nx * h.ny;
    }
}

// allocate_cells_CPU.cpp:33-125: greedy colouring in cell index order
void greedy_groups(const HostCells& h, std::vector<int>& gsize, std::vector<int>& gcells,
                   std::vector<int>& group_of) {
    std::vector<int> type(h.C, 0);
    group_of.assign(h.C, -1);
    gsize.clear();
    gcells.clear();
    int left = h.C;
    int nb[26];
    while (left > 0) {
        int nm = 0;
        int g = (int)gsize.size();
        for (int i = 0; i < h.C; ++i) {
            if (type[i] != 0) continue;
            type[i] = 1;
            gcells.push_back(i);
            group_of[i] = g;
            ++nm;
            --left;
            host_neigh(h, i, nb);
            for (int j = 0; j < 26; ++j)
                if (type[nb[j]] != 1) type[nb[j]] = -1;
        }
        for (int i = 0; i < h.C; ++i)
            if (type[i] == -1) type[i] = 0;
        gsize.push_back(nm);
    }
}

// Dependency levels of a group-ordered Gauss-Seidel sweep. A cell reads the
// cells in `slots`; its update must follow those of earlier groups and
// precede those of later groups. level(c) = 1 + max level of its earlier-group
// readers; cells of one level never read each other, so one launch per level
// reproduces the sweep exactly.
void make_levels(const HostCells& h, const std::vector<int>& gsize, const std::vector<int>& gcells,
                 const std::vector<int>& group_of, const int* slots, int nslots, bool backward,
                 LevelSet& L, std::vector<int>* level_out = nullptr) {
    std::vector<int> level(h.C, -1);
    int G = (int)gsize.size();
    std::vector<int> goff(G + 1, 0);
    for (int g = 0; g < G; ++g) goff[g + 1] = goff[g] + gsize[g];
    int maxl = -1;
    int nb[26];
    for (int gg = 0; gg < G; ++gg) {
        int g = backward ? G - 1 - gg : gg;
        for (int k = goff[g]; k < goff[g + 1]; ++k) {
            int c = gcells[k];
            host_neigh(h, c, nb);
            int lv = 0;
            for (int s = 0; s < nslots; ++s) {
                int n = nb[slots[s]];
                int gn = group_of[n];
                bool earlier = backward ? gn > g : gn < g;
                if (earlier && level[n] + 1 > lv) lv = level[n] + 1;
            }
            level[c] = lv;
            maxl = std::max(maxl, lv);
        }
    }
    if (level_out) *level_out = level;
    L.off.assign(maxl + 2, 0);
    for (int c = 0; c < h.C; ++c) L.off[level[c] + 1]++;
    for (int l = 0; l <= maxl; ++l) L.off[l + 1] += L.off[l];
    L.cells.assign(h.C, 0);
    std::vector<int> fill(L.off.begin(), L.off.end() - 1);
    // keep group order inside a level
    for (int gg = 0; gg < G; ++gg) {
        int g = backward ? G - 1 - gg : gg;
        for (int k = goff[g]; k < goff[g + 1]; ++k) {
            int c = gcells[k];
            L.cells[fill[level[c]]++] = c;
        }
    }
}

// ---------------------------------------------------------------------------
// device building blocks

int reduce(sld_ctx* c, int nblocks, int nk, double* out, size_t part_offset = 0) {
    LAUNCH(c, k_reduce, 1, 256, (const double*)(c->d_part + part_offset), nblocks, nk, out);
    return SLD_OK;
}

// per-replica sums of block partials into each replica's scalar block
int reduce_seg(sld_ctx* c, const int* d_seg, int nk, int out_off) {
    LAUNCH(c, k_reduce_seg, c->R, 1024, (const double*)c->d_part, d_seg, nk, c->d_scal, out_off);
    return SLD_OK;
}

// brings the device noise-pass counter up to the host's
int sync_pass(sld_ctx* c) {
    if (c->pass != c->pass_dev) {
        LAUNCH(c, k_pass_advance, 1, 64, c->D, (int)(c->pass - c->pass_dev));
        c->pass_dev = c->pass;
    }
    return SLD_OK;
}

// the Landau coefficients follow the density
int refresh_landau(sld_ctx* c) {
    LAUNCH(c, k_landau, nblk(c->D.N, 256), 256, c->tab, c->D);
    return SLD_OK;
}

int nbuild(sld_ctx* c, int write_rho) {
    const int w = c->shapes.nbuild_cells;
    LAUNCH(c, k_nbuild, xcd_grid(nblk(c->D.C, w)), 64 * w, c->tab, c->phys, c->D, write_rho);
    c->nb_dirty = false;
    if (write_rho) TRY(refresh_landau(c));
    return SLD_OK;
}

int ensure_nb(sld_ctx* c) {
    if (c->nb_dirty) return nbuild(c, 0);
    return SLD_OK;
}

int do_links(sld_ctx* c) {
    LAUNCH(c, k_bin, nblk(c->D.N, 256), 256, c->D);
    // swap list buffers: the old lists become the input of the rebuild
    std::swap(c->D.cl, c->d_cl2);
    std::swap(c->D.cn, c->d_cn2);
    LAUNCH(c, k_relink, nblk(c->D.C, 128), 128, (const int*)c->d_cl2, (const int*)c->d_cn2, c->D);
    return SLD_OK;
}

int do_force(sld_ctx* c) {
    TRY(ensure_nb(c));
    LAUNCH(c, k_fp, xcd_grid(nblk(c->D.N, 256)), 256, c->tab, c->D);
    LAUNCH(c, k_force, xcd_grid(nblk(c->D.N, c->shapes.force_block)), c->shapes.force_block, c->tab, c->phys,
           c->D);
    if (c->phys.localvol) {
        // calculate_force_energy_CPU.cpp:270-293: the volumes normalised per
        // replica, then the stress over them
        LAUNCH(c, k_lvol_sum, nblk(c->D.N, 256), 256, c->D, c->d_part);
        TRY(reduce_seg(c, c->d_atom_seg, 1, kScalLvol));
        LAUNCH(c, k_lvol_apply, nblk(c->D.N, 256), 256, c->D, (const double*)c->d_scal);
    }
    return SLD_OK;
}

int do_ds(sld_ctx* c, double dt) {
    TRY(ensure_nb(c));
    double dth = dt / 2.0;
    if (c->phys.magmom) LAUNCH(c, k_magmom_s, nblk(c->D.N, 256), 256, c->D, 0, c->D.N);
    for (int dir = 0; dir < 2; ++dir) {
        LevelSet& L = dir == 0 ? c->spin_fwd : c->spin_bwd;
        if (c->phys.spinlang)
            LAUNCH(c, k_spin_noise, nblk(c->D.N, 256), 256, c->D, (int)(c->pass - c->pass_dev));
        for (int l = 0; l < L.nlev(); ++l) {
            const int n = L.off[l + 1] - L.off[l];
            const int bs = c->shapes.spin_block, lpc = c->shapes.spin_lanes;
            const int* cells = L.d_cells + L.off[l];
            // only the backward sweep's Heff_H, Heff_L and sum J|s_j| last
            if (lpc == 4)
                LAUNCH(c, k_spin_phase<4>, nblk(n * 4, bs), bs, cells, n, dir, dir, dth, c->tab, c->phys, c->D);
            else if (lpc == 2)
                LAUNCH(c, k_spin_phase<2>, nblk(n * 2, bs), bs, cells, n, dir, dir, dth, c->tab, c->phys, c->D);
            else
                LAUNCH(c, k_spin_phase<1>, nblk(n, bs), bs, cells, n, dir, dir, dth, c->tab, c->phys, c->D);
        }
        c->pass++;
    }
    if (c->phys.magmom) LAUNCH(c, k_magmom_m, nblk(c->D.N, 256), 256, c->D);
    return SLD_OK;
}

// scalar block of each replica: [0..2] momentum sum, [3..4] kinetic energy
// before and after centring
int do_dp(sld_ctx* c, double dt) {
    int nb = nblk(c->D.C, 8);
    // quantum_noise (quantum_noise_CPU.cpp:60-81): every M-th core_dp_A call
    // redraws the noise of all degrees of freedom at once
    if (c->phys.quantumnoise && c->phys.lattlang) {
        if (c->qcalls % (uint64_t)c->phys.qm == 0)
            LAUNCH(c, k_qn_refresh, nblk(c->D.N, 256), 256, c->phys, c->D, c->qcalls / (uint64_t)c->phys.qm);
        ++c->qcalls;
    }
    // eight cells (8 lanes each) per 64-thread block
    LAUNCH(c, k_dp_cells, nb, 64, dt, c->phys, c->D, (int)(c->pass - c->pass_dev), c->d_part);
    if (c->phys.lattlang) c->pass++;
    TRY(reduce_seg(c, c->d_cell8_seg, 3, 0));
    int na = nblk(c->D.N, 256);
    LAUNCH(c, k_dp_center, na, 256, c->phys, c->D, (const double*)c->d_scal, c->d_part);
    TRY(reduce_seg(c, c->d_atom_seg, 2, 3));
    LAUNCH(c, k_dp_scale, na, 256, c->phys, c->D, (const double*)c->d_scal);
    return SLD_OK;
}

int do_dr(sld_ctx* c, double dt) {
    LAUNCH(c, k_dr, nblk(c->D.N, 256), 256, dt, c->phys, c->D);
    c->nb_dirty = true;
    return SLD_OK;
}

int do_temperature(sld_ctx* c) {
    TRY(ensure_nb(c));
    const int bs = c->shapes.temperature_block;
    LAUNCH(c, k_temperature, xcd_grid(nblk(c->D.C, bs / kSeg)), bs, c->tab, c->phys, c->D);
    return SLD_OK;
}

// core_dTe_CPU.cpp:47-217 with nloop = 50
int do_dTe_launches(sld_ctx* c, double dt);

// core_dTe_CPU.cpp:47-217 through a graph of its launches when graphs are on
// and no kernel is being timed
int do_dTe(sld_ctx* c, double dt) {
    if (c->graph_mode < 1 || !c->timed_name.empty() || c->capturing || c->prm.changestep)
        return do_dTe_launches(c, dt);
    if (c->dte_exec && c->dte_exec_dt != dt) {
        (void)hipGraphExecDestroy(c->dte_exec);
        c->dte_exec = nullptr;
    }
    if (!c->dte_exec) {
        hipGraph_t g = nullptr;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = do_dTe_launches(c, dt);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc != SLD_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return set_err(c, SLD_ERR_HIP, "capture of core_dTe: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&c->dte_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return set_err(c, SLD_ERR_HIP, "instantiate core_dTe graph: %s", hipGetErrorString(e));
        c->dte_exec_dt = dt;
    }
    HIPCHK(c, hipGraphLaunch(c->dte_exec, c->stream));
    return SLD_OK;
}

// first level of a sweep's trailing run of levels that have at most `cap`
// cells each, worth running in one workgroup (two levels or more, at most
// kTailMaxLev); nlev when there is none or cap is 0
int tail_start(const LevelSet& L, int cap) {
    const int nl = L.nlev();
    if (cap <= 0) return nl;
    int lt = nl;
    while (lt > 0 && nl - lt < kTailMaxLev && L.off[lt] - L.off[lt - 1] <= cap) --lt;
    return nl - lt >= 2 ? lt : nl;
}

int do_dTe_launches(sld_ctx* c, double dt) {
    const int nloop = 50;
    double dtl = dt / nloop;
    double h = dtl / 2.0;
    // the electron subsystem runs on single-system contexts only
    const ReplicaHost& h0 = c->rh[0];
    const double natom = c->reps[0].natom, ncell = c->reps[0].ncell;
    double kappa = c->prm.kappa_e / kEvToJ / 1e10;
    double atomic_volume = h0.vol / natom;
    double vf = atomic_volume / (natom / ncell);
    double pref = dtl / 2.0 * kappa * vf / kBoltz;
    int nb = nblk(c->D.C, 256);
    const Mesh M{c->D.Te, c->D.Ee, c->D.Gel, c->D.Ges, c->D.Tl, c->D.TsL, c->D.C, c->phys.lattlang,
                 c->phys.spinlang};
    LAUNCH(c, k_dte_ab, nb, 256, h, M, c->d_part);
    // block partials: [0, nb) starting energies (k_dte_ab / k_dte_finish_ab),
    // then the final energies of each backward level in turn
    for (int it = 0; it < nloop; ++it) {
        int npb = 0;
        // the sweep levels are a few microseconds each: time them as one span
        hipEvent_t* ev = timing_slot(c, "k_dte_sweep");
        if (ev) {
            (void)hipEventRecord(ev[0], c->stream);
            c->ev_span = true;
        }
        int nlaunch = 0;
        bool reduced = false;
        const Box& bx = c->reps[0].box;
        const double ix2 = 1.0 / h0.dx2, iy2 = 1.0 / h0.dy2, iz2 = 1.0 / h0.dz2;
        for (int dir = 0; dir < 2; ++dir) {
            LevelSet& L = dir == 0 ? c->dte_fwd : c->dte_bwd;
            const int lt = tail_start(L, c->shapes.sweep_tail);
            for (int l = 0; l < lt; ++l) {
                int n = L.off[l + 1] - L.off[l];
                double* part = dir == 1 ? c->d_part + (size_t)nb + npb : nullptr;
                const int bs = c->shapes.sweep_block;
                LAUNCH(c, k_dte_sweep, nblk(n, bs), bs, L.d_cells + L.off[l], n, pref, ix2, iy2, iz2, bx.nx,
                       bx.ny, bx.nz, c->D.Te, c->D.Ee, part);
                if (dir == 1) npb += nblk(n, 64);  // one partial per wave
                nlaunch += n > 0;
            }
            if (lt < L.nlev()) {
                // the small trailing levels in one workgroup; after the
                // backward sweep's, the same block adds the sub-step's totals
                SweepTail T{};
                T.nlev = L.nlev() - lt;
                for (int l = lt; l <= L.nlev(); ++l) T.off[l - lt] = L.off[l];
                for (int l = lt; l < L.nlev(); ++l) {
                    T.poff[l - lt] = npb;
                    if (dir == 1) npb += nblk(L.off[l + 1] - L.off[l], 64);
                }
                double* part = dir == 1 ? c->d_part + (size_t)nb : nullptr;
                LAUNCH(c, k_dte_tail, 1, 1024, L.d_cells, T, pref, ix2, iy2, iz2, bx.nx, bx.ny, bx.nz, c->D.Te,
                       c->D.Ee, part, (const double*)c->d_part, nb, (const double*)(c->d_part + nb),
                       dir == 1 ? npb : 0, dir == 1 ? c->d_scal + 8 : nullptr);
                reduced = dir == 1;
                nlaunch += 1;
            }
        }
        if (ev) {
            c->ev_span = false;
            (void)hipEventRecord(ev[1], c->stream);
            c->ev_launches[(ev - c->ev_pool.data()) / 2] = nlaunch;
        }
        if (!reduced)
            LAUNCH(c, k_reduce2, 1, 1024, (const double*)c->d_part, nb, (const double*)(c->d_part + nb), npb,
                   c->d_scal + 8);
        if (it + 1 < nloop)
            LAUNCH(c, k_dte_finish_ab, nb, 256, h, M, (const double*)(c->d_scal + 8), 1, c->d_part);
        else
            LAUNCH(c, k_dte_finish, nb, 256, h, M, (const double*)(c->d_scal + 8), 1);
    }
    return SLD_OK;
}

}  // namespace

// observables record from the reduced sums; renormalizeEnergy error
// (check_energy_CPU.cpp:33-111, check_temperature_CPU.cpp:33-134,
// check_spin_CPU.cpp:33-56)
// One thread per replica; its scalar block holds the atom sums at [16..27]
// and the cell sums at [32..37]. renormalizeEnergy (single system) keeps the
// initial energy and the error in einit[0..1].
__global__ void k_obs_final(const double* scal, const Rep* rep, int R, int eltemp, int renorm, int step,
                            double total_time, double* einit, sld_obs* out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const double* sa = scal + (size_t)r * kScal + 16;
    const double* sc = scal + (size_t)r * kScal + 36;
    const int N = rep[r].natom, C = rep[r].ncell;
    sld_obs o = {};
    o.step = step;
    o.total_time = total_time;
    o.ave_pe = sa[0] / N;
    o.ave_ke = sa[1] / N;
    o.ave_me = sa[2] / N;
    o.ave_me0 = sa[3] / N;
    double err = 0.0;
    if (eltemp) {
        double apc = (double)N / (double)C;
        o.ave_ee = apc * sc[0] / N;
        if (renorm) {
            double e = 0.0;
            e += o.ave_pe; e += o.ave_ke; e += o.ave_me; e += o.ave_me0; e += o.ave_ee;
            if (step == -1) einit[0] = e;
            else err = e - einit[0];
        }
    }
    einit[1] = err;
    o.tl = sa[1] * 2.0 / 3.0 / N / kBoltz;
    double rup = eltemp ? sc[1] : sa[7], rdn = eltemp ? sc[2] : sa[8];
    double lup = eltemp ? sc[3] : sa[9], ldn = eltemp ? sc[4] : sa[10];
    o.ts_r = rup / rdn / 2.0 / kBoltz;
    o.ts_l = lup / ldn / kBoltz;
    if (eltemp) o.te = sc[5] / C / kBoltz;
    double sx = sa[4] / N, sy = sa[5] / N, sz = sa[6] / N;
    o.ave_s[0] = sx; o.ave_s[1] = sy; o.ave_s[2] = sz;
    o.ave_s0 = sqrt(sx * sx + sy * sy + sz * sz);
    o.ave_m[0] = -kElG * sx; o.ave_m[1] = -kElG * sy; o.ave_m[2] = -kElG * sz;
    o.ave_m0 = sqrt(o.ave_m[0] * o.ave_m[0] + o.ave_m[1] * o.ave_m[1] + o.ave_m[2] * o.ave_m[2]);
    o.virial = sa[11];
    // check_stress_CPU.cpp:53-58 and check_pressure_CPU.cpp:35-41 (eV/A^3 -> GPa)
    for (int k = 0; k < 6; ++k) o.ave_stress[k] = sa[12 + k] * (160.217653e0 / N);
    const double density = (double)N / rep[r].box_volume;
    const double tmp = 2e0 / 3e0 * sa[1] / N;
    o.pressure = density * (tmp - o.virial / (3e0 * N));
    o.pressure *= 160.217653e0;
    out[r] = o;
}

__global__ void k_obs_set_te(const double* te_sum, int C, sld_obs* out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) out->te = te_sum[0] / C / kBoltz;
}

namespace {

// check() of the reference (check.cpp:31-47). With renormalizeEnergy the
// reference's energy check also rewrites every cell's Ee and Te, at step -1
// too (Te = T(E(Te)), check_energy_CPU.cpp:100-111); feedback = false computes
// the observables only (sld_create's initial record, which the reference's
// initialize() does not make).
int do_check(sld_ctx* c, int step, sld_obs* d_out, bool feedback = true) {
    int na = nblk(c->D.N, 256);
    int scan = c->phys.eltemp ? 0 : 1;
    if (scan) TRY(ensure_nb(c));
    LAUNCH(c, k_obs_atoms, na, 256, c->tab, c->phys, c->D, c->d_part, scan);
    TRY(reduce_seg(c, c->d_atom_seg, kObsAtomK + 1, 16));
    if (c->phys.eltemp) {
        int nb = nblk(c->D.C, 256);
        LAUNCH(c, k_obs_cells, nb, 256, c->D, c->d_part);
        TRY(reduce(c, nb, kObsCellK, c->d_scal + 36));
    }
    LAUNCH(c, k_obs_final, nblk(c->R, 64), 64, (const double*)c->d_scal, (const Rep*)c->D.rep, c->R,
           c->phys.eltemp, c->prm.renormalize_energy, step, c->total_time, c->d_einit, d_out);
    if (feedback && c->phys.eltemp && c->prm.renormalize_energy) {
        int nb = nblk(c->D.C, 256);
        LAUNCH(c, k_renorm, nb, 256, c->D, (const double*)(c->d_einit + 1), c->d_part);
        TRY(reduce(c, nb, 1, c->d_scal + 44));
        LAUNCH(c, k_obs_set_te, 1, 64, (const double*)(c->d_scal + 44), c->D.C, d_out);
    }
    return SLD_OK;
}

// squared length below which min_image shifts nothing: fractional coordinate
// i of r is row i of the inverse box matrix dotted with r, so |q_i| <= |r|
// times that row's norm; half over the largest norm, less a margin far above
// the rounding of the products, keeps every |q_i| below 1/2
double safe_image_sq(const Box& b) {
    const double n0 = std::sqrt(b.inv[0] * b.inv[0] + b.inv[1] * b.inv[1] + b.inv[3] * b.inv[3]);
    const double n1 = std::sqrt(b.inv[2] * b.inv[2] + b.inv[4] * b.inv[4]);
    const double n2 = std::fabs(b.inv[5]);
    const double r = 0.5 / std::max(n0, std::max(n1, n2)) * (1.0 - 1e-9);
    return r * r;
}

// the geometry a box matrix determines (vec_utils_CPU.cpp:135-148,
// core_dTe_CPU.cpp:143-147): inverse, volume, cell edges; the link-cell
// counts stay
void set_box(sld_ctx* c, int r, const double* d) {
    ReplicaHost& h = c->rh[r];
    Rep& R = c->reps[r];
    Box& b = R.box;
    for (int k = 0; k < 6; ++k) {
        b.d[k] = d[k];
        h.prm.box[k] = d[k];
    }
    b.inv[0] = 1e0 / d[0];
    b.inv[1] = -d[1] / (d[0] * d[2]);
    b.inv[2] = 1e0 / d[2];
    b.inv[3] = (d[1] * d[4] - d[2] * d[3]) / (d[0] * d[2] * d[5]);
    b.inv[4] = -d[4] / (d[2] * d[5]);
    b.inv[5] = 1e0 / d[5];
    b.safe_sq = safe_image_sq(b);
    h.vol = d[0] * d[2] * d[5];
    const double blx = std::fabs(d[0]);
    const double bly = std::sqrt(d[1] * d[1] + d[2] * d[2]);
    const double blz = std::sqrt(d[3] * d[3] + d[4] * d[4] + d[5] * d[5]);
    h.dx2 = std::pow(blx / h.hc.nx, 2);
    h.dy2 = std::pow(bly / h.hc.ny, 2);
    h.dz2 = std::pow(blz / h.hc.nz, 2);
    R.box_volume = h.vol;
    if (r == 0)
        for (int k = 0; k < 6; ++k) c->prm.box[k] = d[k];
}

int upload_reps(sld_ctx* c) {
    HIPCHK(c, hipMemcpyAsync((void*)c->D.rep, c->reps.data(), c->reps.size() * sizeof(Rep), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SLD_OK;
}

// scale_stress_CPU.cpp:33-86: the box matrix moves toward the target
// stresses, the atoms keep their fractional coordinates
int do_scale_stress(sld_ctx* c) {
    const double delta_time = c->total_time - c->last_t_stress;
    c->last_t_stress = c->total_time;
    const double ftmass = 100e0;
    const double pre_fact = delta_time / c->baro.baro_damping_time / ftmass;
    const double* sv = c->last_obs.ave_stress;  // 11 22 33 12 23 31
    const double* t = c->baro.stress;           // xx yx yy zx zy zz
    const double fxx = pre_fact * (sv[0] - t[0]), fyx = pre_fact * (sv[3] - t[1]);
    const double fyy = pre_fact * (sv[1] - t[2]), fzx = pre_fact * (sv[5] - t[3]);
    const double fzy = pre_fact * (sv[4] - t[4]), fzz = pre_fact * (sv[2] - t[5]);
    const Box old_box = c->reps[0].box;
    double d[6];
    for (int k = 0; k < 6; ++k) d[k] = old_box.d[k];
    // in the reference's order: each line sees the entries updated above it
    d[0] = (1e0 + fxx) * d[0];
    d[1] = (1e0 + fxx) * d[1] + fyx * (d[2] + d[0]) / 2e0;
    d[2] = (1e0 + fyy) * d[2];
    d[3] = (1e0 + fxx) * d[3] + fzx * (d[5] + d[0]) / 2e0;
    d[4] = (1e0 + fyy) * d[4] + fzy * (d[5] + d[2]) / 2e0;
    d[5] = (1e0 + fzz) * d[5];
    set_box(c, 0, d);
    LAUNCH(c, k_rebox_r, nblk(c->D.N, 256), 256, c->D, old_box, c->reps[0].box);
    TRY(upload_reps(c));
    c->nb_dirty = true;
    return SLD_OK;
}

// scale_pressure_CPU.cpp:33-61: isotropic scaling toward the target pressure
int do_scale_pressure(sld_ctx* c) {
    const double delta_time = c->total_time - c->last_t_pressure;
    c->last_t_pressure = c->total_time;
    const double ftmass = 100e0;
    const double factor = std::pow(
        1e0 + delta_time / c->baro.baro_damping_time * (c->last_obs.pressure - c->baro.pressure) / ftmass, 1e0 / 3e0);
    double d[6];
    for (int k = 0; k < 6; ++k) d[k] = c->reps[0].box.d[k] * factor;
    set_box(c, 0, d);
    LAUNCH(c, k_scale_r, nblk(c->D.N, 256), 256, c->D, factor);
    TRY(upload_reps(c));
    c->nb_dirty = true;
    return SLD_OK;
}

int do_scale_step(sld_ctx* c);
void drop_graphs(sld_ctx* c);

// scale(step) (scale.cpp:31-44)
int do_scale(sld_ctx* c, int step) {
    bool boxed = false;
    if (c->baro.stress_on && (step + 1) % c->baro.interval_of_scale_stress == 0) {
        TRY(do_scale_stress(c));
        boxed = true;
    }
    if (c->baro.pressure_on && (step + 1) % c->baro.interval_of_scale_pressure == 0) {
        TRY(do_scale_pressure(c));
        boxed = true;
    }
    if (boxed) drop_graphs(c);
    if (c->prm.changestep) TRY(do_scale_step(c));
    return SLD_OK;
}

// scale_step_CPU.cpp:33-97: shrink dt by 0.8 when the step just taken moved
// an atom farther than displace_limit or turned a spin by more than
// phi_limit, else grow it by 1.05
int do_scale_step(sld_ctx* c) {
    const int nb = nblk(c->D.N, 256);
    LAUNCH(c, k_step_max, nb, 256, c->D, c->d_part);
    LAUNCH(c, k_reduce_max2, 1, 256, (const double*)c->d_part, nb, c->d_scal + 48);
    double mx[2] = {0.0, 0.0};
    HIPCHK(c, hipMemcpyAsync(mx, c->d_scal + 48, sizeof(mx), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double displace_max = std::sqrt(mx[0]) / c->prm.atmass * c->prm.step;
    const double phi_max = std::sqrt(mx[1]) / kHbar * c->prm.step;
    const int sw = (displace_max > c->prm.displace_limit ? 1 : 0) + (phi_max > c->prm.phi_limit ? 1 : 0);
    if (sw > 0) c->prm.step *= 0.80;
    else c->prm.step *= 1.05;
    return SLD_OK;
}

// core.cpp:31-76 (SLDHL)
int do_core(sld_ctx* c) {
    double dt = c->prm.step;
    TRY(do_links(c));
    if (c->phys.eltemp) TRY(do_dTe(c, dt / 2.0));
    TRY(do_dp(c, dt / 2.0));
    TRY(do_ds(c, dt / 2.0));
    TRY(do_dr(c, dt));
    TRY(nbuild(c, 1));  // calculate_rho
    TRY(do_ds(c, dt / 2.0));
    TRY(do_force(c));
    TRY(do_dp(c, dt / 2.0));
    if (c->phys.eltemp) {
        TRY(do_temperature(c));
        TRY(do_dTe(c, dt / 2.0));
    }
    return sync_pass(c);
}

void drop_graphs(sld_ctx* c) {
    if (c->dte_exec) (void)hipGraphExecDestroy(c->dte_exec);
    c->dte_exec = nullptr;
    for (auto& g : c->step_exec) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

// one core() step: eager, or (graph_mode 2) through the graph of its launches.
// Capturing runs the host logic of the step (buffer swap, pass counting)
// once; a replay applies the same host effects.
int run_step(sld_ctx* c) {
    if (c->graph_mode < 2 || !c->timed_name.empty() || c->prm.changestep || c->phys.quantumnoise) return do_core(c);
    TRY(ensure_nb(c));  // a step starts with current lists
    const int parity = c->D.cl == c->cl_first ? 0 : 1;
    hipGraphExec_t& ex = c->step_exec[parity];
    if (!ex) {
        const uint64_t p0 = c->pass;
        hipGraph_t g = nullptr;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        c->capturing = true;
        int rc = do_core(c);
        c->capturing = false;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc != SLD_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return set_err(c, SLD_ERR_HIP, "capture of a step: %s", hipGetErrorString(e));
        e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return set_err(c, SLD_ERR_HIP, "instantiate step graph: %s", hipGetErrorString(e));
        c->step_npass = c->pass - p0;
        HIPCHK(c, hipGraphLaunch(ex, c->stream));  // host effects already applied by the capture
        return SLD_OK;
    }
    HIPCHK(c, hipGraphLaunch(ex, c->stream));
    std::swap(c->D.cl, c->d_cl2);
    std::swap(c->D.cn, c->d_cn2);
    c->pass += c->step_npass;
    c->pass_dev = c->pass;
    c->nb_dirty = false;
    return SLD_OK;
}

int read_flags(sld_ctx* c) {
    int flags = 0;
    HIPCHK(c, hipMemcpyAsync(&flags, c->D.err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (flags & kErrCellOverflow)
        return set_err(c, SLD_ERR_CAPACITY, "a link cell exceeded %d atoms", kCellCap);
    if (flags & kErrNbOverflow)
        return set_err(c, SLD_ERR_CAPACITY, "an atom exceeded %d neighbours within rcut_max or %d within rcut_mag",
                       kMaxNb, kMaxMag);
    if (flags & kErrFarMove)
        return set_err(c, SLD_ERR_STATE, "an atom moved beyond the neighbouring link cells in one step");
    if (flags & kErrNonFinite)
        return set_err(c, SLD_ERR_STATE, "non-finite position, momentum or spin");
    return SLD_OK;
}

// sums the event pairs of the timed kernel recorded since ev_used was reset
// (the stream must have passed the last event)
int collect_timing(sld_ctx* c) {
    c->timed_launches = 0;
    c->timed_ms = 0.0;
    for (int k = 0; k + 1 < c->ev_used; k += 2) {
        float kms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&kms, c->ev_pool[k], c->ev_pool[k + 1]));
        c->timed_ms += kms;
        c->timed_launches += c->ev_launches[k / 2];
    }
    return SLD_OK;
}

// host <-> device copies of replica `r`'s atom slots [atom0, atom0 + natom).
// With with_len (spins) the device sets .w = |s| afterwards, with the one
// rounding every other writer of .w uses.
int up4(sld_ctx* c, double4* d, const double* src, int r, bool with_len) {
    if (!src) return SLD_OK;
    const Rep& R = c->reps[r];
    std::vector<double> h((size_t)R.natom * 4);
    for (int i = 0; i < R.natom; ++i) {
        double x = src[3 * i], y = src[3 * i + 1], z = src[3 * i + 2];
        h[4 * i] = x; h[4 * i + 1] = y; h[4 * i + 2] = z;
        h[4 * i + 3] = 0.0;
    }
    HIPCHK(c, hipMemcpyAsync(d + R.atom0, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (with_len) LAUNCH(c, k_spin_len, nblk(R.natom, 256), 256, c->D, R.atom0, R.natom);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SLD_OK;
}

// The caller's spin field of a replica: the spins, or with magmom the moments
// m, from which the spins follow as s = m / (-g) (read_config.cpp:136-139)
int up_spins(sld_ctx* c, const double* src, int r) {
    if (!c->phys.magmom) return up4(c, c->D.s, src, r, true);
    if (!src) return SLD_OK;
    TRY(up4(c, c->D.m, src, r, false));
    const Rep& R = c->reps[r];
    LAUNCH(c, k_magmom_s, nblk(R.natom, 256), 256, c->D, R.atom0, R.natom);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SLD_OK;
}

// electron energy of a cell at Te (eV), in the arithmetic of heatcapacity_CPU.cpp:42-48
double host_te_to_ee(double te) {
    const double bb = 2e-4 / kBoltz;
    return 3e0 / bb * std::log(std::cosh(bb * te));
}

int up1(sld_ctx* c, double* d, const double* src, size_t off, size_t n) {
    if (!src) return SLD_OK;
    HIPCHK(c, hipMemcpyAsync(d + off, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SLD_OK;
}

int down4(sld_ctx* c, const double4* d, double* dst, int r) {
    if (!dst) return SLD_OK;
    const Rep& R = c->reps[r];
    std::vector<double> h((size_t)R.natom * 4);
    HIPCHK(c, hipMemcpyAsync(h.data(), d + R.atom0, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < R.natom; ++i) {
        dst[3 * i] = h[4 * i]; dst[3 * i + 1] = h[4 * i + 1]; dst[3 * i + 2] = h[4 * i + 2];
    }
    return SLD_OK;
}

int down1(sld_ctx* c, const double* d, double* dst, size_t off, size_t n) {
    if (!dst) return SLD_OK;
    HIPCHK(c, hipMemcpyAsync(dst, d + off, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SLD_OK;
}

template <class T>
int upload_vec(sld_ctx* c, T** d, const std::vector<T>& h) {
    TRY(dalloc(c, d, h.size()));
    HIPCHK(c, hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SLD_OK;
}

int upload_levels(sld_ctx* c, LevelSet& L) {
    TRY(upload_vec(c, &L.d_cells, L.cells));
    TRY(upload_vec(c, &L.d_off, L.off));
    return SLD_OK;
}

// The electron sweep reads its cells as packed coordinates ix | iy << 10 |
// iz << 20, which spares the stencil three integer divisions. The electron
// mesh exists on single-system contexts only; batched ones upload plain
// indices, which no sweep reads.
int upload_dte_levels(sld_ctx* c, LevelSet& L) {
    if (c->R == 1) {
        const HostCells& hc = c->rh[0].hc;
        if (hc.nx > 1023 || hc.ny > 1023 || hc.nz > 1023)
            return set_err(c, SLD_ERR_CAPACITY, "the electron mesh needs at most 1023 link cells per side");
        std::vector<int> packed(L.cells.size());
        for (size_t k = 0; k < L.cells.size(); ++k) {
            const int cell = L.cells[k];
            const int ix = cell % hc.nx, iy = (cell / hc.nx) % hc.ny, iz = cell / (hc.nx * hc.ny);
            packed[k] = ix | (iy << 10) | (iz << 20);
        }
        TRY(upload_vec(c, &L.d_cells, packed));
    } else {
        TRY(upload_vec(c, &L.d_cells, L.cells));
    }
    TRY(upload_vec(c, &L.d_off, L.off));
    return SLD_OK;
}

// the level schedules of all replicas as one: level l holds every replica's
// level-l cells (replicas never interact)
void merge_levels(const std::vector<LevelSet>& per, const std::vector<Rep>& reps, LevelSet& out) {
    int nl = 0;
    for (const LevelSet& L : per) nl = std::max(nl, L.nlev());
    out.cells.clear();
    out.off.assign(1, 0);
    for (int l = 0; l < nl; ++l) {
        for (size_t r = 0; r < per.size(); ++r) {
            const LevelSet& L = per[r];
            if (l >= L.nlev()) continue;
            for (int k = L.off[l]; k < L.off[l + 1]; ++k) out.cells.push_back(reps[r].cell0 + L.cells[k]);
        }
        out.off.push_back((int)out.cells.size());
    }
}

inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

// the parameters every replica of a batch must share (one table set, one dt,
// one set of switches)
bool same_shared(const sld_params& a, const sld_params& b) {
    return a.step == b.step && a.atmass == b.atmass && a.gamma_L_over_mass == b.gamma_L_over_mass &&
           a.gamma_S_HL == b.gamma_S_HL && a.kappa_e == b.kappa_e && a.rcut_pot == b.rcut_pot &&
           a.rcut_mag == b.rcut_mag && a.rcut_max == b.rcut_max &&
           a.min_length_link_cell == b.min_length_link_cell && a.start_time == b.start_time &&
           a.interval_of_print_out == b.interval_of_print_out && a.spinlang == b.spinlang &&
           a.lattlang == b.lattlang && a.eltemp == b.eltemp && a.localcolmot == b.localcolmot &&
           a.renormalize_energy == b.renormalize_energy && a.changestep == b.changestep &&
           a.extfield == b.extfield && a.extforce == b.extforce && a.quantumnoise == b.quantumnoise &&
           a.localvol == b.localvol && (!a.localvol || a.rcut_vol == b.rcut_vol) && a.magmom == b.magmom &&
           (!a.quantumnoise ||
            (a.msteps_quantum == b.msteps_quantum && a.nfrequency_quantum == b.nfrequency_quantum));
}

// initial_quantum_noise (quantum_noise_CPU.cpp:83-113): the filter H is the
// cosine transform of H~(w) = sqrt(hbar|w| (1/2 + 1/expm1(hbar|w|/T)))
// * (w h/2) / sin(w h/2) on 2 Nf frequencies up to pi/h, h = M dt / 2,
// with H~(0) = sqrt(T); the same host arithmetic in the same order.
std::vector<double> quantum_filter(const sld_params& P) {
    const int Nf = P.nfrequency_quantum, Nf2 = 2 * Nf;
    const double T = P.temperature;
    const double h = P.msteps_quantum * P.step / 2e0;
    const double pi = 3.141592653589793238462e0;
    const double delta_frequency = pi / h / (double)Nf;
    std::vector<double> Ht(Nf2), H(Nf2, 0.0);
    for (int i = 0; i < Nf2; ++i) {
        const int k = i - Nf;
        const double frequency = delta_frequency * (double)k;
        Ht[i] = std::sqrt(kHbar * std::fabs(frequency) * (0.5e0 + 1e0 / std::expm1(kHbar * std::fabs(frequency) / T)));
        Ht[i] *= (frequency * h / 2e0) / std::sin(frequency * h / 2e0);
        if (k == 0) Ht[i] = std::sqrt(T);
    }
    for (int i = 0; i < Nf2; ++i) {
        for (int j = 0; j < Nf2; ++j) H[i] += Ht[j] * std::cos(pi / (double)Nf * (double)(j - Nf) * (double)(i - Nf));
        H[i] /= (double)Nf2;
    }
    return H;
}

int create_impl(sld_ctx* c, int nrep, const sld_params* params, const sld_atoms* init, int compute_forces) {
    if (nrep <= 0) return set_err(c, SLD_ERR_ARG, "need at least one replica");
    c->R = nrep;
    c->rh.resize(nrep);
    c->reps.resize(nrep);
    for (int r = 0; r < nrep; ++r) {
        sld_params P = params[r];
        const sld_atoms* in = init + r;
        if (P.natom <= 0) return set_err(c, SLD_ERR_ARG, "replica %d: natom must be positive", r);
        if (!in->r || !in->p || !in->s) return set_err(c, SLD_ERR_ARG, "replica %d: r, p and s are required", r);
        if (!(P.box[0] > 0 && P.box[2] > 0 && P.box[5] > 0))
            return set_err(c, SLD_ERR_ARG, "replica %d: box diagonal must be positive", r);
        if (P.rcut_max < P.rcut_pot || P.rcut_max < P.rcut_mag)
            return set_err(c, SLD_ERR_ARG, "rcut_max must be the largest cutoff");
        if (P.rcut_max > P.min_length_link_cell) P.min_length_link_cell = P.rcut_max + 0.001;
        if (P.localcolmot && !P.eltemp) P.localcolmot = 0;
        if (P.renormalize_energy && !P.eltemp) P.renormalize_energy = 0;
        if (P.step <= 0) return set_err(c, SLD_ERR_ARG, "step must be positive");
        if (P.quantumnoise && (P.msteps_quantum < 1 || P.nfrequency_quantum < 1))
            return set_err(c, SLD_ERR_ARG, "quantumnoise needs msteps_quantum >= 1 and nfrequency_quantum >= 1");
        if (r > 0 && !same_shared(P, c->rh[0].prm))
            return set_err(c, SLD_ERR_ARG, "replica %d: dt, mass, damping, cutoffs, print interval and switches "
                                           "must equal replica 0's", r);
        c->rh[r].prm = P;
    }
    c->prm = c->rh[0].prm;
    const sld_params& P0 = c->prm;
    if (nrep > 1 && P0.eltemp)
        return set_err(c, SLD_ERR_ARG, "the electron-temperature model runs on single-system contexts only");
    if (nrep > 1 && P0.changestep)
        return set_err(c, SLD_ERR_ARG, "changestep runs on single-system contexts only (one time step per context)");
    if (P0.localvol && !(P0.rcut_vol > 0.0)) return set_err(c, SLD_ERR_ARG, "localvol needs rcut_vol > 0");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_err(c, SLD_ERR_NOGPU, "no HIP device");
    HIPCHK(c, hipGetDevice(&c->device));
    HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    {
        const char* g = std::getenv("SLD_GRAPHS");
        if (g && g[0] >= '0' && g[0] <= '2') c->graph_mode = g[0] - '0';
    }

    // geometry of every replica (read_config.cpp:110-116, vec_utils_CPU.cpp:135-148,
    // map_cells.cpp:41-60) and its place in the concatenated arrays
    int N = 0, C = 0;
    c->atom_seg.assign(nrep + 1, 0);
    c->cell8_seg.assign(nrep + 1, 0);
    for (int r = 0; r < nrep; ++r) {
        ReplicaHost& h = c->rh[r];
        const double* d = h.prm.box;
        Rep& R = c->reps[r];
        Box& b = R.box;
        for (int k = 0; k < 6; ++k) b.d[k] = d[k];
        b.inv[0] = 1e0 / d[0];
        b.inv[1] = -d[1] / (d[0] * d[2]);
        b.inv[2] = 1e0 / d[2];
        b.inv[3] = (d[1] * d[4] - d[2] * d[3]) / (d[0] * d[2] * d[5]);
        b.inv[4] = -d[4] / (d[2] * d[5]);
        b.inv[5] = 1e0 / d[5];
        b.safe_sq = safe_image_sq(b);
        h.hc = cell_counts(d, h.prm.min_length_link_cell);
        if (h.hc.nx < 3 || h.hc.ny < 3 || h.hc.nz < 3)
            return set_err(c, SLD_ERR_ARG, "replica %d: need at least 3 link cells per side (got %d %d %d)", r,
                           h.hc.nx, h.hc.ny, h.hc.nz);
        b.nx = h.hc.nx; b.ny = h.hc.ny; b.nz = h.hc.nz;
        h.vol = d[0] * d[2] * d[5];
        const double blx = std::fabs(d[0]);
        const double bly = std::sqrt(d[1] * d[1] + d[2] * d[2]);
        const double blz = std::sqrt(d[3] * d[3] + d[4] * d[4] + d[5] * d[5]);
        h.dx2 = std::pow(blx / h.hc.nx, 2);
        h.dy2 = std::pow(bly / h.hc.ny, 2);
        h.dz2 = std::pow(blz / h.hc.nz, 2);
        R.temperature = h.prm.temperature;
        R.box_volume = h.vol;
        R.seed = h.prm.random_seed;
        R.replica = h.prm.replica;
        R.natom = h.prm.natom;
        R.atom0 = round_up(N, 256);
        R.ncell = h.hc.C;
        R.cell0 = round_up(C, 8);
        N = R.atom0 + R.natom;
        C = R.cell0 + R.ncell;
        c->atom_seg[r] = R.atom0 / 256;
        c->cell8_seg[r] = R.cell0 / 8;
    }
    c->atom_seg[nrep] = nblk(N, 256);
    c->cell8_seg[nrep] = nblk(C, 8);

    Phys& ph = c->phys;
    ph.atmass = P0.atmass;
    ph.gamma_L = P0.gamma_L_over_mass * P0.atmass;
    ph.gamma_S = P0.gamma_S_HL;
    ph.rcut_pot = P0.rcut_pot;
    ph.rcut_mag = P0.rcut_mag;
    ph.rcut_pot_sq = std::pow(P0.rcut_pot, 2);
    ph.rcut_mag_sq = std::pow(P0.rcut_mag, 2);
    ph.rcut_max_sq = std::pow(P0.rcut_max, 2);
    ph.spinlang = P0.spinlang;
    ph.lattlang = P0.lattlang;
    ph.eltemp = P0.eltemp;
    ph.localcolmot = P0.localcolmot;
    ph.extfield = P0.extfield;
    ph.extforce = P0.extforce;
    ph.quantumnoise = P0.quantumnoise;
    ph.qm = P0.msteps_quantum;
    ph.qnf2 = 2 * P0.nfrequency_quantum;
    ph.localvol = P0.localvol;
    ph.magmom = P0.magmom;
    ph.rcut_vol = P0.rcut_vol;

    // tables (shared)
    std::shared_ptr<const HostTables> ht = shared_tables(P0.rcut_max, P0.magmom != 0);
    TRY(upload_vec(c, &c->d_prow, ht->prow));
    TRY(upload_vec(c, &c->d_rhotab, ht->rhotab));
    c->tab.prow = c->d_prow;
    c->tab.rhotab = c->d_rhotab;
    c->tab.in_over_rmax = ht->in_over_rmax;
    c->tab.in_over_rhomax = ht->in_over_rhomax;

    // device arrays
    Dev& D = c->D;
    D.N = N;
    D.C = C;
    {
        std::vector<int> arep(N, -1), crep(C, -1);
        for (int r = 0; r < nrep; ++r) {
            const Rep& R = c->reps[r];
            std::fill(arep.begin() + R.atom0, arep.begin() + R.atom0 + R.natom, r);
            std::fill(crep.begin() + R.cell0, crep.begin() + R.cell0 + R.ncell, r);
        }
        Rep* d_rep = nullptr;
        int *d_arep = nullptr, *d_crep = nullptr;
        TRY(upload_vec(c, &d_rep, c->reps));
        TRY(upload_vec(c, &d_arep, arep));
        TRY(upload_vec(c, &d_crep, crep));
        D.rep = d_rep;
        D.arep = d_arep;
        D.crep = d_crep;
        TRY(upload_vec(c, &c->d_atom_seg, c->atom_seg));
        TRY(upload_vec(c, &c->d_cell8_seg, c->cell8_seg));
    }
    TRY(dalloc(c, &D.r, N)); TRY(dalloc(c, &D.p, N)); TRY(dalloc(c, &D.f, N));
    TRY(dalloc(c, &D.s, N)); TRY(dalloc(c, &D.hh, N)); TRY(dalloc(c, &D.hl, N));
    TRY(dalloc(c, &D.rho, N)); TRY(dalloc(c, &D.pe, N)); TRY(dalloc(c, &D.ke, N));
    TRY(dalloc(c, &D.lan, N));
    TRY(dalloc(c, &D.me, N)); TRY(dalloc(c, &D.me0, N)); TRY(dalloc(c, &D.vir, N));
    TRY(dalloc(c, &D.zn, N));
    if (P0.extfield) TRY(dalloc(c, &D.hext, N));  // zero: the reference's default field
    if (P0.extforce) TRY(dalloc(c, &D.fext, N));
    if (P0.quantumnoise) {
        const size_t nf2 = (size_t)ph.qnf2;
        TRY(dalloc(c, &D.qring, nf2 * 3 * (size_t)N));
        TRY(dalloc(c, &D.qval, 3 * (size_t)N));
        std::vector<double> H;
        for (int r = 0; r < nrep; ++r) {
            const std::vector<double> h = quantum_filter(c->rh[r].prm);
            H.insert(H.end(), h.begin(), h.end());
        }
        double* d_qH = nullptr;
        TRY(upload_vec(c, &d_qH, H));
        D.qH = d_qH;
        LAUNCH(c, k_qn_init, nblk(N, 256), 256, ph, D);
    }
    TRY(dalloc(c, &D.pass_base, 1));
    TRY(dalloc(c, &D.sjs, N)); TRY(dalloc(c, &D.fp, N)); TRY(dalloc(c, &D.stress, (size_t)6 * N));
    TRY(dalloc(c, &D.lvol, P0.localvol ? N : 1));
    TRY(dalloc(c, &D.m, P0.magmom ? N : 1));
    TRY(dalloc(c, &D.cell, N)); TRY(dalloc(c, &D.cell_prev, N)); TRY(dalloc(c, &D.cslot, N));
    TRY(dalloc(c, &D.cl, (size_t)C * kCellCap)); TRY(dalloc(c, &D.cn, C));
    TRY(dalloc(c, &D.inc, (size_t)C * kCellCap)); TRY(dalloc(c, &D.inc_n, C));
    TRY(dalloc(c, &c->d_cl2, (size_t)C * kCellCap)); TRY(dalloc(c, &c->d_cn2, C));
    TRY(dalloc(c, &D.nb, (size_t)N * kMaxNb)); TRY(dalloc(c, &D.nbn, N));
    TRY(dalloc(c, &D.nbm, (size_t)N * kMaxMag)); TRY(dalloc(c, &D.nbmn, N));
    TRY(dalloc(c, &D.nbJ, (size_t)N * kMaxMag));
    TRY(dalloc(c, &D.Te, C)); TRY(dalloc(c, &D.Ee, C)); TRY(dalloc(c, &D.Tl, C));
    TRY(dalloc(c, &D.Gel, C)); TRY(dalloc(c, &D.Ges, C)); TRY(dalloc(c, &D.TsR, C));
    TRY(dalloc(c, &D.TsL, C)); TRY(dalloc(c, &D.sRup, C)); TRY(dalloc(c, &D.sRdn, C));
    TRY(dalloc(c, &D.sLup, C)); TRY(dalloc(c, &D.sLdn, C)); TRY(dalloc(c, &D.cellv, C));
    TRY(dalloc(c, &D.err, 1));
    int maxblk = std::max(nblk(N, 256), nblk(C, 8)) + 1;
    TRY(dalloc(c, &c->d_part, (size_t)maxblk * 16));
    TRY(dalloc(c, &c->d_scal, (size_t)kScal * nrep));
    TRY(dalloc(c, &c->d_einit, 4));
    c->obs_cap = 1;
    TRY(dalloc(c, &c->d_obs, nrep));

    // atoms of every replica: positions wrapped as initial_links does
    // (links_CPU.cpp:37-50); initial lists with each atom pushed at the head in
    // ascending index, so a list holds its atoms in descending index (:53-81)
    std::vector<int> cell(N, 0), cn(C, 0), cl((size_t)C * kCellCap, 0), cslot(N, 0);
    std::vector<LevelSet> sf(nrep), sb(nrep), tf(nrep), tb(nrep);
    for (int r = 0; r < nrep; ++r) {
        const Rep& R = c->reps[r];
        ReplicaHost& h = c->rh[r];
        const sld_atoms* in = init + r;
        const Box& b = R.box;
        const double* d = h.prm.box;
        const HostCells& hc = h.hc;
        const int n = R.natom;
        std::vector<double> r3(in->r, in->r + 3 * (size_t)n);
        for (int i = 0; i < n; ++i) {
            double x = r3[3 * i], y = r3[3 * i + 1], z = r3[3 * i + 2];
            double q[3] = {b.inv[0] * x + b.inv[1] * y + b.inv[3] * z, b.inv[2] * y + b.inv[4] * z, b.inv[5] * z};
            int idx = 0;
            for (int k = 0; k < 3; ++k)
                if (q[k] < 0e0 || q[k] >= 1e0) { q[k] -= (int)q[k]; ++idx; }
            if (idx > 0) {
                r3[3 * i] = d[0] * q[0] + d[1] * q[1] + d[3] * q[2];
                r3[3 * i + 1] = d[2] * q[1] + d[4] * q[2];
                r3[3 * i + 2] = d[5] * q[2];
            }
            x = r3[3 * i]; y = r3[3 * i + 1]; z = r3[3 * i + 2];
            double qx = b.inv[0] * x + b.inv[1] * y + b.inv[3] * z, qy = b.inv[2] * y + b.inv[4] * z,
                   qz = b.inv[5] * z;
            int lx = (int)(qx * hc.nx), ly = (int)(qy * hc.ny), lz = (int)(qz * hc.nz);
            lx = std::min(std::max(lx, 0), hc.nx - 1);
            ly = std::min(std::max(ly, 0), hc.ny - 1);
            lz = std::min(std::max(lz, 0), hc.nz - 1);
            cell[R.atom0 + i] = R.cell0 + lx + ly * hc.nx + lz * hc.nx * hc.ny;
        }
        TRY(up4(c, D.r, r3.data(), r, false));
        TRY(up4(c, D.p, in->p, r, false));
        TRY(up4(c, D.f, in->f, r, false));
        TRY(up_spins(c, in->s, r));
        TRY(up4(c, D.hh, in->heff_h, r, false));
        TRY(up4(c, D.hl, in->heff_l, r, false));
        TRY(up1(c, D.rho, in->rho, R.atom0, n));
        TRY(up1(c, D.pe, in->pe, R.atom0, n));
        TRY(up1(c, D.ke, in->ke, R.atom0, n));
        TRY(up1(c, D.me, in->me, R.atom0, n));
        TRY(up1(c, D.me0, in->me0, R.atom0, n));
        if (in->stress)
            for (int k = 0; k < 6; ++k) {
                std::vector<double> st(n);
                for (int i = 0; i < n; ++i) st[i] = in->stress[6 * i + k];
                TRY(up1(c, D.stress, st.data(), (size_t)k * N + R.atom0, n));
            }
        for (int i = n - 1; i >= 0; --i) {
            const int a = R.atom0 + i, cc = cell[a];
            if (cn[cc] >= kCellCap)
                return set_err(c, SLD_ERR_CAPACITY, "replica %d: a link cell exceeds %d atoms at start", r, kCellCap);
            cl[(size_t)cc * kCellCap + cn[cc]++] = a;
        }
        // groups and sweep levels (allocate_cells_CPU.cpp:33-125)
        std::vector<int> group_of;
        greedy_groups(hc, h.group_size, h.group_cells, group_of);
        int all26[26];
        for (int k = 0; k < 26; ++k) all26[k] = k;
        const int faces[6] = {0, 13, 2, 14, 12, 15};
        make_levels(hc, h.group_size, h.group_cells, group_of, all26, 26, false, sf[r]);
        make_levels(hc, h.group_size, h.group_cells, group_of, all26, 26, true, sb[r]);
        make_levels(hc, h.group_size, h.group_cells, group_of, faces, 6, false, tf[r]);
        make_levels(hc, h.group_size, h.group_cells, group_of, faces, 6, true, tb[r]);
    }
    for (int cc = 0; cc < C; ++cc)
        for (int q = 0; q < cn[cc]; ++q) cslot[cl[(size_t)cc * kCellCap + q]] = cc * kCellCap + q;
    HIPCHK(c, hipMemcpyAsync(D.cell, cell.data(), N * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(D.cell_prev, cell.data(), N * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(D.cn, cn.data(), C * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(D.cl, cl.data(), cl.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(D.cslot, cslot.data(), N * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    merge_levels(sf, c->reps, c->spin_fwd);
    merge_levels(sb, c->reps, c->spin_bwd);
    merge_levels(tf, c->reps, c->dte_fwd);
    merge_levels(tb, c->reps, c->dte_bwd);
    // A wave of the spin sweep walks as many atoms as its fullest cell holds,
    // and the cells of one level are independent, so each level lists its
    // cells by descending starting atom count (stable: neighbouring cells stay
    // together within a count): waves then hold cells of like length and the
    // longest start first. The order of a level does not change any result.
    for (LevelSet* L : {&c->spin_fwd, &c->spin_bwd})
        for (int l = 0; l < L->nlev(); ++l)
            std::stable_sort(L->cells.begin() + L->off[l], L->cells.begin() + L->off[l + 1],
                             [&](int a, int b) { return cn[a] > cn[b]; });
    TRY(upload_levels(c, c->spin_fwd));
    TRY(upload_levels(c, c->spin_bwd));
    TRY(upload_dte_levels(c, c->dte_fwd));
    TRY(upload_dte_levels(c, c->dte_bwd));

    // electron temperatures (map_cells.cpp:128-159)
    if (P0.eltemp) {
        std::vector<double> te(C, P0.temperature), ee(C);
        for (int k = 0; k < C; ++k) ee[k] = host_te_to_ee(te[k]);
        TRY(up1(c, D.Te, te.data(), 0, C));
        TRY(up1(c, D.Ee, ee.data(), 0, C));
    }
    // initial.cpp:78-87 (with the densities of the input when they are not computed)
    TRY(nbuild(c, compute_forces ? 1 : 0));
    if (!compute_forces) TRY(refresh_landau(c));
    if (compute_forces) TRY(do_force(c));
    if (P0.eltemp) TRY(do_temperature(c));
    c->total_time = P0.start_time;
    c->pass = 0;
    c->pass_dev = 0;
    c->cl_first = c->D.cl;
    c->last_t_stress = c->last_t_pressure = c->total_time;  // initial.cpp:93-98
    TRY(do_check(c, -1, c->d_obs, false));
    HIPCHK(c, hipMemcpyAsync(&c->last_obs, c->d_obs, sizeof(sld_obs), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return read_flags(c);
}

}  // namespace

// ===========================================================================
// C ABI

extern "C" {

int sld_version(void) { return SLD_ABI_VERSION; }

int sld_plan(const double* box, double min_length_link_cell, int* ncell3, int* ngroups,
             int* group_sizes, int* group_cells, int* nlevels, int* levels) {
    if (!box || !ncell3) return SLD_ERR_ARG;
    HostCells hc = cell_counts(box, min_length_link_cell);
    ncell3[0] = hc.nx; ncell3[1] = hc.ny; ncell3[2] = hc.nz;
    if (!group_sizes) return SLD_OK;
    if (hc.nx < 3 || hc.ny < 3 || hc.nz < 3) return SLD_ERR_ARG;
    std::vector<int> gs, gc, gof;
    greedy_groups(hc, gs, gc, gof);
    if (ngroups) *ngroups = (int)gs.size();
    std::copy(gs.begin(), gs.end(), group_sizes);
    if (group_cells) std::copy(gc.begin(), gc.end(), group_cells);
    int all26[26];
    for (int k = 0; k < 26; ++k) all26[k] = k;
    const int faces[6] = {0, 13, 2, 14, 12, 15};
    for (int sch = 0; sch < 4; ++sch) {
        LevelSet L;
        std::vector<int> lv;
        bool bwd = sch % 2 == 1;
        if (sch < 2) make_levels(hc, gs, gc, gof, all26, 26, bwd, L, &lv);
        else make_levels(hc, gs, gc, gof, faces, 6, bwd, L, &lv);
        if (nlevels) nlevels[sch] = L.nlev();
        if (levels) std::copy(lv.begin(), lv.end(), levels + (size_t)sch * hc.C);
    }
    return SLD_OK;
}

int sld_device_info(char* name, int len, int* n_cu) {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess) return SLD_ERR_NOGPU;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return SLD_ERR_NOGPU;
    if (name && len > 0) {
        snprintf(name, len, "%s", prop.gcnArchName);
    }
    if (n_cu) *n_cu = prop.multiProcessorCount;
    return SLD_OK;
}

int sld_create(const sld_params* params, const sld_atoms* init, int compute_initial_forces, sld_ctx** out) {
    return sld_create_batch(1, params, init, compute_initial_forces, out);
}

int sld_create_batch(int nrep, const sld_params* params, const sld_atoms* init, int compute_initial_forces,
                     sld_ctx** out) {
    if (!params || !init || !out) return SLD_ERR_ARG;
    *out = nullptr;
    sld_ctx* c = new sld_ctx();
    int rc = create_impl(c, nrep, params, init, compute_initial_forces);
    *out = c;  // returned even on failure so that sld_last_error can be read
    return rc;
}

int sld_nreplicas(const sld_ctx* c) { return c ? c->R : 0; }

int sld_select_replica(sld_ctx* c, int r) {
    if (!c) return SLD_ERR_ARG;
    if (r < 0 || r >= c->R) return set_err(c, SLD_ERR_ARG, "replica %d out of range (%d replicas)", r, c->R);
    c->sel = r;
    return SLD_OK;
}

int sld_destroy(sld_ctx* c) {
    if (!c) return SLD_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SLD_OK;
}

const char* sld_last_error(const sld_ctx* c) { return c ? c->err.c_str() : "null context"; }

int sld_links(sld_ctx* c) { if (!c) return SLD_ERR_ARG; TRY(do_links(c)); return read_flags(c); }
int sld_calculate_rho(sld_ctx* c) { if (!c) return SLD_ERR_ARG; TRY(nbuild(c, 1)); return read_flags(c); }
int sld_calculate_force_energy(sld_ctx* c) { if (!c) return SLD_ERR_ARG; TRY(do_force(c)); return read_flags(c); }
int sld_core_ds(sld_ctx* c, double dt) {
    if (!c) return SLD_ERR_ARG;
    TRY(do_ds(c, dt));
    TRY(sync_pass(c));
    return read_flags(c);
}
int sld_core_dp(sld_ctx* c, double dt) {
    if (!c) return SLD_ERR_ARG;
    TRY(do_dp(c, dt));
    TRY(sync_pass(c));
    return read_flags(c);
}
int sld_core_dr(sld_ctx* c, double dt) { if (!c) return SLD_ERR_ARG; TRY(do_dr(c, dt)); return read_flags(c); }

int sld_core_dTe(sld_ctx* c, double dt) {
    if (!c->phys.eltemp) return set_err(c, SLD_ERR_STATE, "core_dTe needs eltemp");
    c->ev_used = 0;
    TRY(do_dTe(c, dt));
    TRY(read_flags(c));
    return collect_timing(c);
}

int sld_calculate_temperature(sld_ctx* c) {
    if (!c->phys.eltemp) return set_err(c, SLD_ERR_STATE, "calculate_temperature needs eltemp");
    TRY(do_temperature(c));
    return read_flags(c);
}

int sld_check(sld_ctx* c, int step, sld_obs* out) {
    if (!c) return SLD_ERR_ARG;
    TRY(do_check(c, step, c->d_obs));
    if (out) HIPCHK(c, hipMemcpyAsync(out, c->d_obs, sizeof(sld_obs) * c->R, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&c->last_obs, c->d_obs, sizeof(sld_obs), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return read_flags(c);
}

int sld_step(sld_ctx* c, int first_step, int nsteps, sld_obs* obs, int* nobs) {
    if (!c) return SLD_ERR_ARG;
    if (nsteps < 0) return set_err(c, SLD_ERR_ARG, "nsteps < 0");
    int every = c->prm.interval_of_print_out > 0 ? c->prm.interval_of_print_out : 1;
    int nprint = 0;
    for (int s = first_step; s < first_step + nsteps; ++s)
        if ((s + 1) % every == 0) ++nprint;
    if (nprint > c->obs_cap) {
        sld_obs* nb = nullptr;
        HIPCHK(c, hipMalloc(&nb, sizeof(sld_obs) * nprint * c->R));
        c->allocs.push_back(nb);
        c->alloc_bytes.push_back(sizeof(sld_obs) * nprint * c->R);
        c->d_obs = nb;
        c->obs_cap = nprint;
    }
    c->ev_used = 0;
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    HIPCHK(c, hipEventRecord(e0, c->stream));
    int k = 0;
    for (int s = first_step; s < first_step + nsteps; ++s) {
        TRY(run_step(c));
        c->total_time += c->prm.step;
        if ((s + 1) % every == 0) {
            TRY(do_check(c, s, c->d_obs + (size_t)k * c->R));
            if (c->baro.stress_on || c->baro.pressure_on) {
                HIPCHK(c, hipMemcpyAsync(&c->last_obs, c->d_obs + (size_t)k * c->R, sizeof(sld_obs),
                                         hipMemcpyDeviceToHost, c->stream));
                HIPCHK(c, hipStreamSynchronize(c->stream));
            }
            ++k;
        }
        // scale() (spilady.cpp:50-57, scale.cpp:31-44)
        TRY(do_scale(c, s));
    }
    LAUNCH(c, k_check_nonfinite, nblk(c->D.N, 256), 256, c->D);
    HIPCHK(c, hipEventRecord(e1, c->stream));
    HIPCHK(c, hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
    c->last_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    TRY(collect_timing(c));
    if (obs && k > 0)
        HIPCHK(c, hipMemcpy(obs, c->d_obs, sizeof(sld_obs) * k * c->R, hipMemcpyDeviceToHost));
    if (nobs) *nobs = k;
    return read_flags(c);
}

int sld_last_step_timing(sld_ctx* c, double* total_ms) {
    if (!c) return SLD_ERR_ARG;
    if (total_ms) *total_ms = c->last_ms;
    return SLD_OK;
}

int sld_time_kernel(sld_ctx* c, const char* kernel) {
    if (!c) return SLD_ERR_ARG;
    c->timed_name = kernel ? kernel : "";
    c->ev_used = 0;
    return SLD_OK;
}

int sld_kernel_timing(sld_ctx* c, int* launches, double* total_ms) {
    if (!c) return SLD_ERR_ARG;
    if (launches) *launches = c->timed_launches;
    if (total_ms) *total_ms = c->timed_ms;
    return SLD_OK;
}

int sld_download(sld_ctx* c, sld_atoms* o) {
    if (!c) return SLD_ERR_ARG;
    if (!o) return SLD_ERR_ARG;
    Dev& D = c->D;
    const int r = c->sel;
    const Rep& R = c->reps[r];
    const size_t a0 = R.atom0, n = R.natom;
    TRY(down4(c, D.r, o->r, r)); TRY(down4(c, D.p, o->p, r)); TRY(down4(c, D.f, o->f, r));
    TRY(down4(c, c->phys.magmom ? D.m : D.s, o->s, r));
    TRY(down4(c, D.hh, o->heff_h, r)); TRY(down4(c, D.hl, o->heff_l, r));
    TRY(down1(c, D.rho, o->rho, a0, n)); TRY(down1(c, D.pe, o->pe, a0, n)); TRY(down1(c, D.ke, o->ke, a0, n));
    TRY(down1(c, D.me, o->me, a0, n)); TRY(down1(c, D.me0, o->me0, a0, n)); TRY(down1(c, D.vir, o->vir, a0, n));
    TRY(down1(c, D.sjs, o->sum_jij_sj, a0, n));
    if (o->local_volume) {
        if (c->phys.localvol) {
            TRY(down1(c, D.lvol, o->local_volume, a0, n));
        } else {
            // calculate_force_energy_CPU.cpp:276
            for (size_t i = 0; i < n; ++i) o->local_volume[i] = R.box_volume / R.natom;
        }
    }
    if (o->stress) {
        std::vector<double> st(n);
        for (int k = 0; k < 6; ++k) {
            TRY(down1(c, D.stress, st.data(), (size_t)k * D.N + a0, n));
            for (size_t i = 0; i < n; ++i) o->stress[6 * i + k] = st[i];
        }
    }
    return SLD_OK;
}

int sld_upload(sld_ctx* c, const sld_atoms* in) {
    if (!c) return SLD_ERR_ARG;
    if (!in) return SLD_ERR_ARG;
    Dev& D = c->D;
    const int r = c->sel;
    const Rep& R = c->reps[r];
    TRY(up4(c, D.r, in->r, r, false));
    TRY(up4(c, D.p, in->p, r, false));
    TRY(up_spins(c, in->s, r));
    TRY(up4(c, D.f, in->f, r, false));
    TRY(up1(c, D.rho, in->rho, R.atom0, R.natom));
    TRY(refresh_landau(c));
    if (in->r) c->nb_dirty = true;
    return SLD_OK;
}

int sld_get_cells(sld_ctx* c, sld_cells* o) {
    if (!c) return SLD_ERR_ARG;
    if (!o) return SLD_ERR_ARG;
    Dev& D = c->D;
    const Rep& R = c->reps[c->sel];
    const size_t c0 = R.cell0, n = R.ncell;
    TRY(down1(c, D.Te, o->te, c0, n)); TRY(down1(c, D.Ee, o->ee, c0, n)); TRY(down1(c, D.Tl, o->tl, c0, n));
    TRY(down1(c, D.TsR, o->ts_r, c0, n)); TRY(down1(c, D.TsL, o->ts_l, c0, n));
    TRY(down1(c, D.Gel, o->gel, c0, n)); TRY(down1(c, D.Ges, o->ges, c0, n));
    TRY(down1(c, D.sRup, o->sum_r_up, c0, n)); TRY(down1(c, D.sRdn, o->sum_r_dn, c0, n));
    TRY(down1(c, D.sLup, o->sum_l_up, c0, n)); TRY(down1(c, D.sLdn, o->sum_l_dn, c0, n));
    return SLD_OK;
}

int sld_set_cells_te(sld_ctx* c, const double* te) {
    if (!c || !te) return SLD_ERR_ARG;
    const Rep& R = c->reps[c->sel];
    // Ee = E(Te) on the host, as map_cells.cpp:154 and sld_create compute it
    std::vector<double> ee(R.ncell);
    for (int k = 0; k < R.ncell; ++k) ee[k] = host_te_to_ee(te[k]);
    TRY(up1(c, c->D.Te, te, R.cell0, R.ncell));
    TRY(up1(c, c->D.Ee, ee.data(), R.cell0, R.ncell));
    return SLD_OK;
}

int sld_get_geometry(sld_ctx* c, int* ncell3, int* ngroups, int* group_sizes) {
    if (!c) return SLD_ERR_ARG;
    const ReplicaHost& h = c->rh[c->sel];
    if (ncell3) { ncell3[0] = h.hc.nx; ncell3[1] = h.hc.ny; ncell3[2] = h.hc.nz; }
    if (ngroups) *ngroups = (int)h.group_size.size();
    if (group_sizes) std::copy(h.group_size.begin(), h.group_size.end(), group_sizes);
    return SLD_OK;
}

int sld_get_cell_lists(sld_ctx* c, int cap, int* counts, int* lists) {
    if (!c) return SLD_ERR_ARG;
    const Rep& R = c->reps[c->sel];
    const int C = R.ncell;
    std::vector<int> cn(C), cl((size_t)C * kCellCap);
    HIPCHK(c, hipMemcpyAsync(cn.data(), c->D.cn + R.cell0, C * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(cl.data(), c->D.cl + (size_t)R.cell0 * kCellCap, cl.size() * sizeof(int),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int mx = 0;
    for (int k = 0; k < C; ++k) {
        if (counts) counts[k] = cn[k];
        mx = std::max(mx, cn[k]);
        if (lists)
            for (int q = 0; q < std::min(cn[k], cap); ++q)
                lists[(size_t)k * cap + q] = cl[(size_t)k * kCellCap + q] - R.atom0;
    }
    if (mx > cap) return set_err(c, SLD_ERR_CAPACITY, "cell holds %d atoms > cap %d", mx, cap);
    return SLD_OK;
}

double sld_total_time(const sld_ctx* c) { if (!c) return 0.0; return c->total_time; }

int sld_scale_step(sld_ctx* c) {
    if (!c) return SLD_ERR_ARG;
    TRY(do_scale_step(c));
    return read_flags(c);
}

double sld_get_step(const sld_ctx* c) { return c ? c->prm.step : 0.0; }

int sld_set_barostat(sld_ctx* c, const sld_barostat* b) {
    if (!c || !b) return SLD_ERR_ARG;
    if ((b->stress_on || b->pressure_on) && c->R > 1)
        return set_err(c, SLD_ERR_ARG, "barostats run on single-system contexts only");
    if (b->stress_on && b->interval_of_scale_stress < 1) return set_err(c, SLD_ERR_ARG, "stress interval < 1");
    if (b->pressure_on && b->interval_of_scale_pressure < 1) return set_err(c, SLD_ERR_ARG, "pressure interval < 1");
    if ((b->stress_on || b->pressure_on) && !(b->baro_damping_time > 0.0))
        return set_err(c, SLD_ERR_ARG, "baro_damping_time must be positive");
    c->baro = *b;
    return SLD_OK;
}

int sld_scale(sld_ctx* c, int step) {
    if (!c) return SLD_ERR_ARG;
    TRY(do_scale(c, step));
    return read_flags(c);
}

int sld_scale_stress(sld_ctx* c) {
    if (!c) return SLD_ERR_ARG;
    if (!c->baro.stress_on) return set_err(c, SLD_ERR_STATE, "the STRESS barostat is off (sld_set_barostat)");
    TRY(do_scale_stress(c));
    drop_graphs(c);
    return read_flags(c);
}

int sld_scale_pressure(sld_ctx* c) {
    if (!c) return SLD_ERR_ARG;
    if (!c->baro.pressure_on) return set_err(c, SLD_ERR_STATE, "the PRESSURE barostat is off (sld_set_barostat)");
    TRY(do_scale_pressure(c));
    drop_graphs(c);
    return read_flags(c);
}

int sld_set_total_time(sld_ctx* c, double t) {
    if (!c) return SLD_ERR_ARG;
    if (!std::isfinite(t)) return set_err(c, SLD_ERR_ARG, "total_time is not finite");
    c->total_time = t;
    return SLD_OK;
}

int sld_get_box(sld_ctx* c, double* box) {
    if (!c || !box) return SLD_ERR_ARG;
    for (int k = 0; k < 6; ++k) box[k] = c->reps[c->sel].box.d[k];
    return SLD_OK;
}

int sld_set_external(sld_ctx* c, const double* hext, const double* fext) {
    if (!c) return SLD_ERR_ARG;
    if (hext && !c->phys.extfield) return set_err(c, SLD_ERR_STATE, "extfield is off in sld_params");
    if (fext && !c->phys.extforce) return set_err(c, SLD_ERR_STATE, "extforce is off in sld_params");
    TRY(up4(c, c->D.hext, hext, c->sel, false));
    TRY(up4(c, c->D.fext, fext, c->sel, false));
    return SLD_OK;
}

// ---------------------------------------------------------------------------
// launch shapes and the autotuner

}  // extern "C"

namespace {
constexpr int kNShape = 7;
const int kShapeCand[kNShape][3] = {{32, 64, 128}, {64, 128, 256}, {64, 128, 256}, {1, 2, 4}, {64, 128, 256},
                                    {0, 2048, 8192}, {1, 2, 4}};
// the kernel each shape is timed on (the sweep span covers the tail launches)
const char* kShapeKernel[kNShape] = {"k_spin_phase", "k_dte_sweep", "k_force", "k_nbuild", "k_temperature",
                                     "k_dte_sweep", "k_spin_phase"};

int& shape_field(sld_shapes& s, int k) {
    switch (k) {
        case 0: return s.spin_block;
        case 1: return s.sweep_block;
        case 2: return s.force_block;
        case 3: return s.nbuild_cells;
        case 4: return s.temperature_block;
        case 5: return s.sweep_tail;
        default: return s.spin_lanes;
    }
}

bool shape_valid(const sld_shapes& s) {
    sld_shapes t = s;
    for (int k = 0; k < kNShape; ++k) {
        const int v = shape_field(t, k);
        if (v != kShapeCand[k][0] && v != kShapeCand[k][1] && v != kShapeCand[k][2]) return false;
    }
    return true;
}


// a device copy of every allocation and the host fields that a step changes
struct Snapshot {
    void* buf = nullptr;
    Dev D{};
    int *cl2 = nullptr, *cn2 = nullptr;
    uint64_t pass = 0, pass_dev = 0, qcalls = 0;
    double total_time = 0.0;
    bool nb_dirty = true;
};

int snapshot_take(sld_ctx* c, Snapshot& s) {
    size_t tot = 0;
    for (size_t b : c->alloc_bytes) tot += (b + 255) & ~size_t(255);
    HIPCHK(c, hipMalloc(&s.buf, tot));
    size_t off = 0;
    for (size_t k = 0; k < c->allocs.size(); ++k) {
        HIPCHK(c, hipMemcpyAsync((char*)s.buf + off, c->allocs[k], c->alloc_bytes[k], hipMemcpyDeviceToDevice,
                                 c->stream));
        off += (c->alloc_bytes[k] + 255) & ~size_t(255);
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    s.D = c->D;
    s.cl2 = c->d_cl2;
    s.cn2 = c->d_cn2;
    s.pass = c->pass;
    s.pass_dev = c->pass_dev;
    s.qcalls = c->qcalls;
    s.total_time = c->total_time;
    s.nb_dirty = c->nb_dirty;
    return SLD_OK;
}

int snapshot_restore(sld_ctx* c, Snapshot& s) {
    if (!s.buf) return SLD_OK;
    (void)hipStreamSynchronize(c->stream);
    size_t off = 0;
    for (size_t k = 0; k < c->allocs.size(); ++k) {
        HIPCHK(c, hipMemcpyAsync(c->allocs[k], (char*)s.buf + off, c->alloc_bytes[k], hipMemcpyDeviceToDevice,
                                 c->stream));
        off += (c->alloc_bytes[k] + 255) & ~size_t(255);
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipFree(s.buf));
    s.buf = nullptr;
    c->D = s.D;
    c->d_cl2 = s.cl2;
    c->d_cn2 = s.cn2;
    c->pass = s.pass;
    c->pass_dev = s.pass_dev;
    c->qcalls = s.qcalls;
    c->total_time = s.total_time;
    c->nb_dirty = s.nb_dirty;
    return SLD_OK;
}

// device microseconds per step spent in kernel `name` over `steps` steps
int time_in_steps(sld_ctx* c, const char* name, int steps, double* us, int* launches) {
    c->timed_name = name;
    c->ev_used = 0;
    for (int k = 0; k < steps; ++k) TRY(do_core(c));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    TRY(collect_timing(c));
    c->timed_name.clear();
    *us = c->timed_ms * 1e3 / steps;
    *launches = c->timed_launches;
    return SLD_OK;
}

int autotune_body(sld_ctx* c, int steps, double* us) {
    TRY(do_core(c));  // warm-up
    for (int k = 0; k < kNShape; ++k) {
        double best = 0.0;
        int best_v = shape_field(c->shapes, k);
        for (int q = 0; q < 3; ++q) {
            shape_field(c->shapes, k) = kShapeCand[k][q];
            double t = 0.0;
            int n = 0;
            TRY(time_in_steps(c, kShapeKernel[k], steps, &t, &n));
            if (us) us[3 * k + q] = n > 0 ? t : 0.0;
            if (n > 0 && (q == 0 || t < best)) {
                best = t;
                best_v = kShapeCand[k][q];
            }
        }
        shape_field(c->shapes, k) = best_v;
    }
    return read_flags(c);
}
}  // namespace

extern "C" {

int sld_set_graph_mode(sld_ctx* c, int mode) {
    if (!c) return SLD_ERR_ARG;
    if (mode < 0 || mode > 2) return set_err(c, SLD_ERR_ARG, "graph mode %d is not 0, 1 or 2", mode);
    if (mode != c->graph_mode) drop_graphs(c);
    c->graph_mode = mode;
    return SLD_OK;
}

int sld_get_shapes(sld_ctx* c, sld_shapes* out) {
    if (!c || !out) return SLD_ERR_ARG;
    *out = c->shapes;
    return SLD_OK;
}

int sld_set_shapes(sld_ctx* c, const sld_shapes* in) {
    if (!c || !in) return SLD_ERR_ARG;
    if (!shape_valid(*in)) return set_err(c, SLD_ERR_ARG, "launch shape outside the supported candidates");
    c->shapes = *in;
    drop_graphs(c);
    return SLD_OK;
}

int sld_autotune(sld_ctx* c, int steps, sld_shapes* best, double* us) {
    if (!c) return SLD_ERR_ARG;
    if (steps < 1) return set_err(c, SLD_ERR_ARG, "autotune needs at least one timed step");
    if (us)
        for (int k = 0; k < 3 * kNShape; ++k) us[k] = 0.0;
    const sld_shapes before = c->shapes;
    Snapshot snap;
    TRY(snapshot_take(c, snap));
    drop_graphs(c);
    int rc = autotune_body(c, steps, us);
    c->timed_name.clear();
    const std::string msg = c->err;
    int rc2 = snapshot_restore(c, snap);
    drop_graphs(c);
    if (rc != SLD_OK) {
        c->shapes = before;
        c->err = msg;
        return rc;
    }
    if (rc2 != SLD_OK) return rc2;
    if (best) *best = c->shapes;
    return SLD_OK;
}

// ---------------------------------------------------------------------------
// parametric potential model (sld_model.h, sld_model.hip)

int sld_model_nparams(void) { return kNParam; }

int sld_model_default_params(double* theta) {
    if (!theta) return SLD_ERR_ARG;
    model_default_params(theta);
    return SLD_OK;
}

int sld_model_set_params(sld_ctx* c, const double* theta) {
    if (!c) return SLD_ERR_ARG;
    if (!theta) return set_err(c, SLD_ERR_ARG, "null parameter vector");
    for (int k = 0; k < kNParam; ++k)
        if (!std::isfinite(theta[k])) return set_err(c, SLD_ERR_ARG, "parameter %d is not finite", k);
    if (!(theta[P_JRC] > 0.0)) return set_err(c, SLD_ERR_ARG, "exchange cutoff must be positive");
    auto& m = c->model;
    const int N = c->D.N;
    if (!m.d_r) {
        TRY(dalloc(c, &m.d_r, (size_t)kNInput * kMRStride));
        TRY(dalloc(c, &m.d_rd, (size_t)kNInput * kMRDStride));
        TRY(dalloc(c, &m.d_h, (size_t)kNInput * kMHStride));
        std::vector<int> coff(c->R);
        int tot = 0;
        for (int r = 0; r < c->R; ++r) {
            coff[r] = tot;
            tot += c->reps[r].natom;
        }
        m.ntot = tot;
        TRY(upload_vec(c, &m.d_coff, coff));
        TRY(dalloc(c, &m.d_rho, N)); TRY(dalloc(c, &m.d_fp, N)); TRY(dalloc(c, &m.d_Q, N));
        TRY(dalloc(c, &m.d_F, (size_t)3 * tot)); TRY(dalloc(c, &m.d_H, (size_t)3 * tot));
        TRY(dalloc(c, &m.d_gF, (size_t)3 * tot)); TRY(dalloc(c, &m.d_gH, (size_t)3 * tot));
        TRY(dalloc(c, &m.d_gE, c->R)); TRY(dalloc(c, &m.d_E, (size_t)kScal * c->R));
        TRY(dalloc(c, &m.d_grad, kNParam));
        TRY(dalloc(c, &m.d_part, (size_t)nblk(N, 256) * kNParam));
    }
    ModelHostTables t;
    build_model_tables(theta, c->prm.rcut_max, t);
    HIPCHK(c, hipMemcpyAsync(m.d_r, t.r.data(), t.r.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(m.d_rd, t.rd.data(), t.rd.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(m.d_h, t.h.data(), t.h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    m.tab.r = m.d_r;
    m.tab.rd = m.d_rd;
    m.tab.h = m.d_h;
    m.tab.in_over_rmax = t.in_over_rmax;
    m.tab.in_over_rhomax = t.in_over_rhomax;
    for (int k = 0; k < 4; ++k) m.tab.L[k] = theta[P_LA + k];
    std::copy(theta, theta + kNParam, m.theta);
    m.ready = true;
    return SLD_OK;
}

int sld_model_get_params(sld_ctx* c, double* theta) {
    if (!c || !theta) return SLD_ERR_ARG;
    if (!c->model.ready) return set_err(c, SLD_ERR_STATE, "model parameters not set");
    std::copy(c->model.theta, c->model.theta + kNParam, theta);
    return SLD_OK;
}

}  // extern "C"

namespace {
// densities and F'(rho) of the current configurations at the current parameters
int model_rho(sld_ctx* c) {
    auto& m = c->model;
    if (!m.ready) return set_err(c, SLD_ERR_STATE, "model parameters not set (sld_model_set_params)");
    TRY(ensure_nb(c));
    TRY(read_flags(c));
    LAUNCH(c, k_model_rho, xcd_grid(nblk(c->D.N, 256)), 256, m.tab, c->phys, c->D, m.d_rho, m.d_fp);
    return SLD_OK;
}

// copy between a caller's buffer (host or device) and a device buffer
int copy_any(sld_ctx* c, void* dst, const void* src, size_t bytes) {
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, c->stream));
    return SLD_OK;
}
}  // namespace

extern "C" {

int sld_model_eval(sld_ctx* c, double* energy, double* force, double* field) {
    if (!c) return SLD_ERR_ARG;
    auto& m = c->model;
    c->ev_used = 0;
    TRY(model_rho(c));
    const int nb = nblk(c->D.N, 256);
    LAUNCH(c, k_model_forward, xcd_grid(nb), 256, m.tab, c->phys, c->D, (const double*)m.d_rho, (const double*)m.d_fp,
           (const int*)m.d_coff, m.d_F, m.d_H, m.d_part);
    LAUNCH(c, k_reduce_seg, c->R, 1024, (const double*)m.d_part, (const int*)c->d_atom_seg, 1, m.d_E, 0);
    if (energy)
        HIPCHK(c, hipMemcpy2DAsync(energy, sizeof(double), m.d_E, kScal * sizeof(double), sizeof(double), c->R,
                                   hipMemcpyDefault, c->stream));
    const size_t b3 = (size_t)3 * m.ntot * sizeof(double);
    if (force) TRY(copy_any(c, force, m.d_F, b3));
    if (field) TRY(copy_any(c, field, m.d_H, b3));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return collect_timing(c);
}

int sld_model_vjp(sld_ctx* c, const double* g_energy, const double* g_force, const double* g_field,
                  double* grad) {
    if (!c) return SLD_ERR_ARG;
    if (!grad) return set_err(c, SLD_ERR_ARG, "null gradient output");
    auto& m = c->model;
    c->ev_used = 0;
    TRY(model_rho(c));
    const size_t b3 = (size_t)3 * m.ntot * sizeof(double);
    if (g_energy) TRY(copy_any(c, m.d_gE, g_energy, c->R * sizeof(double)));
    else HIPCHK(c, hipMemsetAsync(m.d_gE, 0, c->R * sizeof(double), c->stream));
    if (g_force) TRY(copy_any(c, m.d_gF, g_force, b3));
    else HIPCHK(c, hipMemsetAsync(m.d_gF, 0, b3, c->stream));
    if (g_field) TRY(copy_any(c, m.d_gH, g_field, b3));
    else HIPCHK(c, hipMemsetAsync(m.d_gH, 0, b3, c->stream));
    const int nb = nblk(c->D.N, 256);
    LAUNCH(c, k_model_q, xcd_grid(nb), 256, m.tab, c->phys, c->D, (const int*)m.d_coff, (const double*)m.d_gF, m.d_Q);
    LAUNCH(c, k_model_adjoint, xcd_grid(nb), 256, m.tab, c->phys, c->D, (const double*)m.d_rho, (const double*)m.d_fp,
           (const double*)m.d_Q, (const int*)m.d_coff, (const double*)m.d_gE, (const double*)m.d_gF,
           (const double*)m.d_gH, m.d_part);
    LAUNCH(c, k_reduce, 1, 256, (const double*)m.d_part, nb, kNParam, m.d_grad);
    TRY(copy_any(c, grad, m.d_grad, kNParam * sizeof(double)));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return collect_timing(c);
}

int sld_model_natoms(const sld_ctx* c) {
    if (!c) return 0;
    int tot = 0;
    for (const Rep& r : c->reps) tot += r.natom;
    return tot;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// sld_model_eval_noalloc: sld_model_eval for callers that cannot allocate
// device memory. Its scratch (densities and F'(rho) per atom slot, the block
// partials of the energy and the per-replica sums) lives in the file-scope
// array below, of a fixed size, which every GPU the library runs on holds
// until the process ends. Host code reaches a device's copy only through the
// address hipGetSymbolAddress returns for that device.
namespace {
constexpr int kNoAllocAtoms = 1 << 18;                   // atom slots, padding included
constexpr int kNoAllocBlocks = kNoAllocAtoms / 256 + 8;  // 256-thread blocks of the energy pass
constexpr int kNoAllocReps = 64;                         // replicas
constexpr size_t kNoAllocDoubles =
    2 * (size_t)kNoAllocAtoms + (size_t)kNoAllocBlocks + (size_t)kScal * kNoAllocReps;
constexpr int kMaxDevices = 64;

std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {
// the current device's copy of `scratch`, looked up once per device
int noalloc_scratch(sld_ctx* c, double** out) {
    int dev = 0;
    HIPCHK(c, hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_err(c, SLD_ERR_ARG, "device %d outside the scratch table", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_err(c, SLD_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                           hipGetErrorString(e));
        g_scratch_dev[dev] = (double*)p;
    }
    *out = g_scratch_dev[dev];
    return SLD_OK;
}
}  // namespace

extern "C" {

int sld_model_eval_noalloc(sld_ctx* c, double* energy, double* force, double* field) {
    if (!c) return SLD_ERR_ARG;
    auto& m = c->model;
    if (!m.ready) return set_err(c, SLD_ERR_STATE, "model parameters not set (sld_model_set_params)");
    const int N = c->D.N, nb = nblk(N, 256);
    if (N > kNoAllocAtoms || nb > kNoAllocBlocks || c->R > kNoAllocReps)
        return set_err(c, SLD_ERR_CAPACITY,
                       "sld_model_eval_noalloc: %d atom slots in %d replicas exceed its fixed scratch "
                       "(%d atom slots, %d replicas); use sld_model_eval",
                       N, c->R, kNoAllocAtoms, kNoAllocReps);
    double* base = nullptr;
    TRY(noalloc_scratch(c, &base));
    double* rho = base;
    double* fp = base + kNoAllocAtoms;
    double* part = base + 2 * (size_t)kNoAllocAtoms;
    double* esum = part + kNoAllocBlocks;
    c->ev_used = 0;
    // every scratch element read below is written earlier in this call: the
    // density pass writes rho and fp of every real atom slot (the only ones
    // the energy pass reads), the energy pass one partial per block, the
    // segmented sum one value per replica
    TRY(ensure_nb(c));
    TRY(read_flags(c));
    LAUNCH(c, k_model_rho, xcd_grid(nb), 256, m.tab, c->phys, c->D, rho, fp);
    double* F = force ? force : m.d_F;
    double* H = field ? field : m.d_H;
    LAUNCH(c, k_model_forward, xcd_grid(nb), 256, m.tab, c->phys, c->D, (const double*)rho, (const double*)fp,
           (const int*)m.d_coff, F, H, part);
    LAUNCH(c, k_reduce_seg, c->R, 1024, (const double*)part, (const int*)c->d_atom_seg, 1, esum, 0);
    if (energy)
        HIPCHK(c, hipMemcpy2DAsync(energy, sizeof(double), esum, kScal * sizeof(double), sizeof(double), c->R,
                                   hipMemcpyDefault, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return collect_timing(c);
}

}  // extern "C"

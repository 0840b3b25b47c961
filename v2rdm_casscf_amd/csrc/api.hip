// api.hip -- the v2g C ABI (include/v2g.h): solver handle, device buffers,
// the CG driver, the Update_xz driver and the BPSDP outer loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/v2g.h"
#include "common.hpp"
#include "devlayout.hpp"
#include "eig.hpp"
#include "inputs.hpp"
#include "kernels.hpp"
#include "layout.hpp"
#include "mlp.hpp"
#include "spatial.hpp"

#include <random>

namespace v2g {
std::string &last_error()
{
    thread_local std::string slot;
    return slot;
}
}  // namespace v2g

namespace {

using v2g::DevBuf;
using v2g::HipError;
using v2g::now;

// Scratch of the allocation-free CG entry point: one array per GPU in the code
// object's static device memory, 64 MiB.  Host code never passes `scratch`
// itself (in host code the name is a stand-in, not a device address): every
// handle resolves the device address once at creation, on its device, and
// passes that pointer to the kernels.
constexpr size_t kScratchDoubles = (size_t(64) << 20) / sizeof(double);
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

// device address of `scratch` on the current device (resolved once per device)
double *scratch_for_current_device()
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw HipError("device index out of range for the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = static_cast<double *>(ptr);
    }
    return g_scratch_dev[dev];
}

constexpr int kCgChunk = 8;                 // CG iterations between host checks (even)
constexpr int64_t kGraphMaxRows = 4 << 20;  // replay captured chunks below this many rows
// 1: a CG iteration takes p.Ap = ||A^T p||^2 from the A^T.p launch and folds
// the residual update into the A.x launch (layouts without second-pass rows)
#ifndef V2G_CG_FUSE
#define V2G_CG_FUSE 1
#endif

}  // namespace

struct v2g_solver {
    std::unique_ptr<v2g::Layout> L;
    hipStream_t own = nullptr, st = nullptr;
    // layout on device
    DevBuf<int> d_sym, d_gab, d_gaa, d_bab, d_baa, d_baab, d_baba, d_baaa, d_b00;
    DevBuf<v2g::DevLayout> d_lay;
    DevBuf<v2g::Segment> d_rowsegs, d_xsegs, d_t2segs;
    DevBuf<int64_t> d_rowtiles, d_xtiles, d_t2tiles;
    DevBuf<int2> d_g2idx;
    v2g::MapPlan plan;
    // CG
    DevBuf<double> r, p, z, Ap, t, partial, scal, jacobi, user_minv, rowpart;
    bool have_jacobi = false, have_user = false;
    // loop scratch
    DevBuf<double> Ax, Bv, ATy, red;
    DevBuf<double> yhist;   // previous outer iteration's CG solution
    DevBuf<double> xprev;   // x on entry to the last in-loop split
    int mlp_outer_trained = 0;  // outer iterations of MLP training done so far
    bool mlp_fallback = false, mlp_first_learned = true;
    int mlp_last_base_iters = 0;   // CG iterations of the last training solve
    int mlp_jacobi_base = 0;       // 1: learned scaling of the Jacobi diagonal
    // A/B guard of the learned diagonal: CG convergence rates ln(r0 / r) per
    // iteration of the last base solves and of the first learned solves
    double base_rate_sum = 0.0, learned_rate_sum = 0.0;
    int base_rate_n = 0, learned_rate_n = 0;
    double last_cg_nrm0 = 0.0;     // ||r|| at the start of the last CG solve
    // eigen split
    std::vector<v2g::EigBlock> small, large;
    DevBuf<v2g::EigBlock> d_small, d_large;
    DevBuf<double> Mw, Aw, Tw, Vw, Qbuf, tol, frop;
    DevBuf<int> active, counter;
    DevBuf<v2g::GemmDesc> d_gemm_warm1, d_gemm_warm2, d_gemm_rec;
    int max_nb = 0, max_N = 0;
    // block-Jacobi lanes: the large blocks are split into up to kMaxLanes groups
    // of equal work (sum of N^3), each advancing its rounds on its own stream,
    // so that latency-bound pair problems of some lanes overlap the
    // bandwidth-bound tile updates of others.  Lane l holds
    // large[lane_off[l] .. lane_off[l] + lane_n[l]); lane 0 runs on the engine
    // stream, lane l > 0 on lane_st[l].  Two lanes: HIP maps a process's
    // streams onto a few hardware queues, and with four lanes two of them
    // shared a queue and ran one after the other (measured at H50).
    static constexpr int kMaxLanes = 2;
    // largest padded block from which a problem is treated as latency-bound:
    // lanes and 1024-thread pair workgroups
    static constexpr int kLaneMinN = 512;
    int nlanes = 1;
    int lane_off[kMaxLanes] = {}, lane_n[kMaxLanes] = {}, lane_pairs[kMaxLanes] = {},
        lane_nb[kMaxLanes] = {};
    int64_t lane_tiles[kMaxLanes] = {};
    hipStream_t lane_st[kMaxLanes] = {};
    hipEvent_t ev_fork = nullptr, ev_start[kMaxLanes] = {}, ev_join[kMaxLanes] = {};
    std::vector<v2g::EigBlock> large_nat;   // the large blocks in dimensions_ order
    // 1024-thread pair workgroups (latency-bound rounds) instead of 256
    bool pair_wide = false;
    bool warm = false;
    int last_sweeps = 0;
    // warm calls whose sweeps were capped and still rotated: the split is an
    // approximation of the eigendecomposition (tracking), not a converged one
    bool last_split_capped = false;
    const double *rec_x = nullptr;          // x the reconstruction descriptors point at
    // autotuner: the warm splits of the first calls try every launch shape
    // (lanes, pair workgroup) kTuneReps times each and keep the fastest per sweep
    static constexpr int kTuneReps = 2;
    bool tuning = false;
    int tune_k = 0;
    std::vector<std::pair<int, int>> tune_cand;   // (lanes, 1024-thread pairs)
    std::vector<float> tune_ms;                  // best ms per sweep per candidate
    hipEvent_t ev_t0 = nullptr, ev_t1 = nullptr;
    v2g_iter_callback cb = nullptr;
    void *cb_user = nullptr;
    double *scratch_dev = nullptr;          // this device's `scratch` (allocation-free CG)
    // learned preconditioner
    DevBuf<double> rownorm;
    bool have_rownorm = false;
    DevBuf<unsigned char> code;
    DevBuf<float> W, mlp_part, mlp_lpart, feat, tgt, wt;
    DevBuf<double> am, av, grad, dloss, scales, yprev, lminv, vtmp, kd, kd2, rhi, ones;
    DevBuf<int64_t> bidx;
    int adam_t = 0, batch_cap = 0;
    bool mlp_ready = false;

    // CG iterations replayed as a captured HIP graph of kCgChunk iterations
    // (problems up to kGraphMaxRows rows, where launch latency dominates).
    // Opt-in (V2G_CG_GRAPHS=1): with many handles driven from concurrent host
    // threads, a capture on one stream has made other threads' copies fail.
    bool use_graphs = false;
    hipGraphExec_t cg_exec = nullptr;
    const double *cg_exec_y = nullptr, *cg_exec_minv = nullptr;
    int cg_exec_flex = -1;

    explicit v2g_solver(const v2g_layout_desc &d);
    ~v2g_solver()
    {
        if (cg_exec) (void)hipGraphExecDestroy(cg_exec);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_t0) (void)hipEventDestroy(ev_t0);
        if (ev_t1) (void)hipEventDestroy(ev_t1);
        for (int l = 0; l < kMaxLanes; l++) {
            if (ev_start[l]) (void)hipEventDestroy(ev_start[l]);
            if (ev_join[l]) (void)hipEventDestroy(ev_join[l]);
            if (lane_st[l]) (void)hipStreamDestroy(lane_st[l]);
        }
        if (own) (void)hipStreamDestroy(own);
    }
    void configure_lanes(int want);
    void set_shape(int lanes, int wide);
    void start_tuning();
    void ensure_jacobi();
    void ensure_mlp();
    void mlp_init(unsigned long long seed);
    void mlp_scales(const double *rhs, const double *y);
    void mlp_infer(const double *rhs, const double *y);
    double mlp_train(const double *rhs, const double *rlo, int nb, int steps, double lr,
                     double pen, std::mt19937_64 &rng);
    // the device vectors and scalars one CG solve works in
    struct CgBufs {
        double *r, *p, *Ap, *t, *partial, *scal, *rowpart;
    };
    CgBufs own_cg_bufs() { return {r.p, p.p, Ap.p, t.p, partial.p, scal.p, rowpart.p}; }
    // fused iterations: rowpart holds x_tiles squares of A^T p, then three
    // arrays of row_tiles residual partials (the unfused p.Ap partials use the
    // first row_tiles)
    bool fused_cg() const { return V2G_CG_FUSE && plan.t2_tiles == 0 && plan.x_tiles > 0; }
    int64_t rowpart_len() const { return std::max<int64_t>(1, plan.x_tiles + 3 * plan.row_tiles); }
    int cg(const double *rhs, double *y, double tol, int maxiter, int precond, int flexible,
           double *res);
    int cg_with(const CgBufs &B, bool graphs, const double *rhs, double *y, double tol, int maxiter,
                const double *minv, int flexible, double *res);
    void update_xz(double mu, const double *y, const double *c, double *x, double *z,
                   int warm_sweep_cap = 0, int inner_sweep_cap = 0, int cross_order = 0);
    double dot(const double *a, const double *b, int64_t n);
};

v2g_solver::v2g_solver(const v2g_layout_desc &d)
{
    L.reset(new v2g::Layout(v2g::options_from_desc(d)));
    if (const char *e = std::getenv("V2G_CG_GRAPHS")) use_graphs = std::atoi(e) != 0;
    HIPCHK(hipStreamCreateWithFlags(&own, hipStreamNonBlocking));
    st = own;
    scratch_dev = scratch_for_current_device();

    d_sym.upload(L->sym.empty() ? std::vector<int>{0} : L->sym);
    d_gab.upload(L->gab.empty() ? std::vector<int>{0} : L->gab);
    d_gaa.upload(L->gaa.empty() ? std::vector<int>{0} : L->gaa);
    d_bab.upload(L->bab);
    d_baa.upload(L->baa);
    d_baab.upload(L->baab);
    d_baba.upload(L->baba);
    d_baaa.upload(L->baaa);
    d_b00.upload(L->b00);
    const v2g::DevLayout dl = v2g::make_dev_layout(
        *L, {d_sym.p, d_gab.p, d_gaa.p, d_bab.p, d_baa.p, d_baab.p, d_baba.p, d_baaa.p, d_b00.p});
    d_lay.upload(std::vector<v2g::DevLayout>{dl});

    auto tiles = [](const std::vector<v2g::Segment> &segs, std::vector<int64_t> &ts) {
        ts.assign(1, 0);
        for (const auto &s : segs) ts.push_back(ts.back() + v2g::tiles_for(s));
        return ts.back();
    };
    std::vector<int64_t> ts;
    plan.dlay = d_lay.p;
    plan.row_tiles = tiles(L->rowsegs, ts);
    // G2 row segments get a slice of the precomputed column table (aux)
    std::vector<v2g::Segment> rsegs = L->rowsegs;
    int64_t g2rows = 0;
    for (auto &s : rsegs) {
        const bool g2 = s.kind == v2g::RF_G2AB || s.kind == v2g::RF_G2AA;
        s.aux = g2 && !s.wave ? g2rows : -1;
        if (s.aux >= 0) g2rows += s.count;
    }
    d_rowsegs.upload(rsegs);
    d_rowtiles.upload(ts);
    plan.rowsegs = d_rowsegs.p;
    plan.row_tile_start = d_rowtiles.p;
    plan.nrowsegs = (int)L->rowsegs.size();
    plan.x_tiles = tiles(L->xsegs, ts);
    d_xsegs.upload(L->xsegs);
    d_xtiles.upload(ts);
    plan.xsegs = d_xsegs.p;
    plan.x_tile_start = d_xtiles.p;
    plan.nxsegs = (int)L->xsegs.size();
    plan.t2_tiles = tiles(L->xsegs_t2, ts);
    d_t2segs.upload(L->xsegs_t2);
    d_t2tiles.upload(ts);
    plan.t2segs = d_t2segs.p;
    plan.t2_tile_start = d_t2tiles.p;
    plan.nt2segs = (int)L->xsegs_t2.size();
    // the D1a / D1b blocks must form one range of x that no other block
    // enters, and every column must fit 30 bits, for the packed G2 table
    int64_t d1lo = INT64_MAX, d1hi = -1;
    for (int hh = 0; hh < L->nirrep; hh++) {
        const int64_t nn = (int64_t)L->opt.amopi[hh] * L->opt.amopi[hh];
        if (!nn) continue;
        d1lo = std::min(d1lo, L->xoff[v2g::XB_D1A][hh]);
        d1hi = std::max(d1hi, L->xoff[v2g::XB_D1B][hh] + nn);
    }
    bool packable = g2rows > 0 && d1hi > d1lo && L->dimx < ((int64_t)1 << 30);
    for (size_t bi = 0; packable && bi < L->block_dims.size(); bi++) {
        const int64_t lo = L->block_offs[bi], hi = lo + (int64_t)L->block_dims[bi] * L->block_dims[bi];
        bool d1 = false;
        for (int hh = 0; hh < L->nirrep; hh++)
            d1 = d1 || lo == L->xoff[v2g::XB_D1A][hh] || lo == L->xoff[v2g::XB_D1B][hh];
        if (!d1 && hi > lo && lo < d1hi && hi > d1lo) packable = false;
    }
    if (packable) {
        d_g2idx.alloc(g2rows);
        plan.g2idx = d_g2idx.p;
        v2g::build_g2_index(plan, rsegs, d1lo, d1hi, d_g2idx.p, st);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
    }

    const int64_t ncon = L->ncon, dimx = L->dimx;
    r.alloc(ncon);
    p.alloc(ncon);

    Ap.alloc(ncon);
    t.alloc(dimx);
    partial.alloc(v2g::kPartialArrays * v2g::kRedBlocks);
    rowpart.alloc(rowpart_len());
    scal.alloc(v2g::CG_NSLOTS);
    Ax.alloc(ncon);
    Bv.alloc(ncon);
    ATy.alloc(dimx);
    red.alloc(4 * v2g::kRedBlocks + 8);   // up to four partial arrays, then their sums

    // eigen-split plan: blocks <= 64 whole in one workgroup, others block Jacobi
    int64_t woff = 0, qoff = 0, poff = 0;
    for (size_t b = 0; b < L->block_dims.size(); b++) {
        const int n = L->block_dims[b];
        if (n == 0) continue;
        v2g::EigBlock e{};
        e.n = n;
        e.xoff = L->block_offs[b];
        if (n <= 64) {
            e.N = 64;
            small.push_back(e);
            continue;
        }
        e.N = ((n + 63) / 64) * 64;
        e.woff = woff;
        e.qoff = qoff;
        e.poff = poff;
        const int np = e.N / 64;
        woff += (int64_t)e.N * e.N;
        qoff += (int64_t)np * 64 * 64;
        poff += np;
        max_nb = std::max(max_nb, e.N / 32);
        max_N = std::max(max_N, e.N);
        large.push_back(e);
    }
    large_nat = large;
    // problem size decides the default launch shape: from a padded block of
    // kLaneMinN on, the rounds are latency-bound (two lanes, 1024-thread pairs)
    pair_wide = max_N >= kLaneMinN;
    configure_lanes(max_N >= kLaneMinN ? kMaxLanes : 1);
    d_small.upload(small);
    if (!large.empty()) {
        Mw.alloc(woff);
        Aw.alloc(woff);
        Tw.alloc(woff);
        Vw.alloc(woff);
        Qbuf.alloc(qoff);
        active.alloc(poff);
        tol.alloc(large.size());
        frop.alloc(64 * large.size());
        counter.alloc(1);
    }
}

// Split the large blocks into `want` lanes of equal work (sum of N^3,
// longest-processing-time order, largest blocks first).  Blocks are
// independent, so every lane count gives the same split bit for bit; the
// kept eigenvectors live at fixed workspace offsets and stay valid.
void v2g_solver::configure_lanes(int want)
{
    if (want < 1 || want > kMaxLanes) throw std::invalid_argument("lanes must be 1 or 2");
    HIPCHK(hipStreamSynchronize(st));
    for (int l = 1; l < kMaxLanes; l++)
        if (lane_st[l]) HIPCHK(hipStreamSynchronize(lane_st[l]));
    nlanes = std::min<int>(want, std::max<int>(1, (int)large_nat.size()));
    std::vector<v2g::EigBlock> sorted = large_nat, g[kMaxLanes];
    std::stable_sort(sorted.begin(), sorted.end(),
                     [](const v2g::EigBlock &x, const v2g::EigBlock &y) { return x.N > y.N; });
    double load[kMaxLanes] = {};
    for (const auto &e : sorted) {
        int l = 0;
        for (int k = 1; k < nlanes; k++)
            if (load[k] < load[l]) l = k;
        g[l].push_back(e);
        load[l] += (double)e.N * e.N * e.N;
    }
    large.clear();
    for (int l = 0; l < kMaxLanes; l++) {
        lane_off[l] = (int)large.size();
        lane_n[l] = l < nlanes ? (int)g[l].size() : 0;
        lane_pairs[l] = lane_nb[l] = 0;
        lane_tiles[l] = 0;
        if (l < nlanes) large.insert(large.end(), g[l].begin(), g[l].end());
        if (l < nlanes && nlanes > 1) {
            if (l > 0 && !lane_st[l]) HIPCHK(hipStreamCreateWithFlags(&lane_st[l], hipStreamNonBlocking));
            if (!ev_start[l]) HIPCHK(hipEventCreateWithFlags(&ev_start[l], hipEventDisableTiming));
            if (!ev_join[l]) HIPCHK(hipEventCreateWithFlags(&ev_join[l], hipEventDisableTiming));
        }
    }
    if (nlanes > 1 && !ev_fork) HIPCHK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    for (int l = 0; l < nlanes; l++)
        for (int i = lane_off[l]; i < lane_off[l] + lane_n[l]; i++) {
            const int np = large[i].N / 64;
            lane_pairs[l] = std::max(lane_pairs[l], np);
            lane_nb[l] = std::max(lane_nb[l], large[i].N / 32);
            lane_tiles[l] = std::max(lane_tiles[l], (int64_t)np * (np + 1) / 2 + (int64_t)np * np);
        }
    d_large.upload(large.empty() ? std::vector<v2g::EigBlock>{} : large);
    // descriptors of the warm-start products follow the block order
    d_gemm_warm1.alloc(0);
    d_gemm_warm2.alloc(0);
    rec_x = nullptr;
}

void v2g_solver::set_shape(int lanes, int wide)
{
    if (lanes > 0 && lanes != nlanes) configure_lanes(lanes);
    if (wide >= 0) pair_wide = wide != 0;
}

void v2g_solver::start_tuning()
{
    tune_cand.clear();
    const int maxl = std::min<int>(kMaxLanes, (int)large_nat.size());
    for (int l = 1; l <= maxl; l++)
        for (int w = 0; w < 2; w++) tune_cand.push_back({l, w});
    tune_ms.assign(tune_cand.size(), 0.0f);
    tune_k = 0;
    tuning = tune_cand.size() > 1;
    if (tuning && !ev_t0) {
        HIPCHK(hipEventCreate(&ev_t0));
        HIPCHK(hipEventCreate(&ev_t1));
    }
}

void v2g_solver::ensure_jacobi()
{
    if (have_jacobi) return;
    rownorm.alloc(L->ncon);
    jacobi.alloc(L->ncon);
    v2g::launch_row_norm2(plan, rownorm.p, st);
    std::vector<double> d(L->ncon);
    HIPCHK(hipMemcpyAsync(d.data(), rownorm.p, sizeof(double) * d.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (auto &v : d) v = v > 0.0 ? 1.0 / v : 1.0;
    HIPCHK(hipMemcpyAsync(jacobi.p, d.data(), sizeof(double) * d.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    have_jacobi = true;
    have_rownorm = true;
}

void v2g_solver::ensure_mlp()
{
    if (mlp_ready) return;
    ensure_jacobi();
    const int64_t n = L->ncon;
    code.alloc(n);
    v2g::mlp_family_codes(d_rowsegs.p, L->rowsegs, code.p, st);
    W.alloc(v2g::kMlpParams);
    am.alloc(v2g::kMlpParams);
    av.alloc(v2g::kMlpParams);
    grad.alloc(v2g::kMlpParams);
    dloss.alloc(1);
    scales.alloc(8);
    yprev.alloc(n);
    lminv.alloc(n);
    vtmp.alloc(n);
    kd.alloc(n);
    kd2.alloc(n);
    rhi.alloc(n);
    mlp_ready = true;
    mlp_init(12345);
}

void v2g_solver::mlp_init(unsigned long long seed)
{
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<float> w(v2g::kMlpParams, 0.0f);
    const int in = v2g::kMlpIn, hd = v2g::kMlpH;
    for (int i = 0; i < in * hd; i++) w[i] = (float)(nd(rng) / std::sqrt((double)in));
    const int oW2 = in * hd + hd;
    for (int i = 0; i < hd * hd; i++) w[oW2 + i] = (float)(nd(rng) / std::sqrt((double)hd));
    // output layer (w3, b3) stays zero: g = 0, the preconditioner starts as its
    // base (the identity, i.e. the reference's plain CG, or Jacobi)
    HIPCHK(hipMemcpyAsync(W.p, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemsetAsync(am.p, 0, sizeof(double) * v2g::kMlpParams, st));
    HIPCHK(hipMemsetAsync(av.p, 0, sizeof(double) * v2g::kMlpParams, st));
    HIPCHK(hipStreamSynchronize(st));
    adam_t = 0;
    mlp_outer_trained = 0;
    mlp_fallback = false;
    mlp_first_learned = true;
    mlp_last_base_iters = 0;
    base_rate_sum = learned_rate_sum = 0.0;
    base_rate_n = learned_rate_n = 0;
}

// scales[0] = 1/rms(rhs), scales[1] = 1/rms(y)
void v2g_solver::mlp_scales(const double *rhs, const double *y)
{
    const double n = (double)L->ncon;
    const double rr = dot(rhs, rhs, L->ncon), yy = dot(y, y, L->ncon);
    double sc[2] = {rr > 0 ? std::sqrt(n / rr) : 1.0, yy > 0 ? std::sqrt(n / yy) : 1.0};
    HIPCHK(hipMemcpyAsync(scales.p, sc, sizeof(sc), hipMemcpyHostToDevice, st));
}

void v2g_solver::mlp_infer(const double *rhs, const double *y)
{
    mlp_scales(rhs, y);
    v2g::mlp_minv(L->ncon, code.p, rownorm.p, rhs, y, W.p, scales.p, lminv.p, 3.0, mlp_jacobi_base,
                  st);
}

// Online training of the learned diagonal S on the spectral objective
// L = log rho(r_hi) - log rho(r_lo), rho(r) = (Sr).K(Sr) / (r.Sr), K = A A^T:
// r_lo is the final residual of the base CG solve just done (its slowest
// modes), r_hi comes from power iterations r <- K S r (the fastest).  L
// estimates log cond(S K); each Adam step lowers it.  Every step: inference
// on all rows, two K products, four dots, a sampled batch with the exact
// upstream gradient, backprop and Adam.  Returns L after the last step minus
// L of the base diagonal on the same probes (negative: the learned diagonal
// is better conditioned on them).
double v2g_solver::mlp_train(const double *rhs, const double *rlo, int nb, int steps, double lr,
                             double pen, std::mt19937_64 &rng)
{
    const int64_t n = L->ncon;
    if (nb > batch_cap) {
        bidx.alloc(nb);
        feat.alloc((size_t)nb * v2g::kMlpIn);
        wt.alloc(nb);
        mlp_part.alloc((size_t)v2g::mlp_grad_blocks(nb) * v2g::kMlpParams);
        mlp_lpart.alloc(v2g::mlp_grad_blocks(nb));
        batch_cap = nb;
    }
    auto kprod = [&](const double *sv, const double *r, double *out) {   // out = K (s o r)
        v2g::vec_mul(n, sv, r, vtmp.p, st);
        v2g::launch_atu(plan, vtmp.p, t.p, st);
        v2g::launch_au(plan, t.p, out, st);
    };
    auto rho_terms = [&](const double *sv, const double *r, double *kdv, double &a, double &bq) {
        kprod(sv, r, kdv);
        a = dot(r, vtmp.p, n);       // r . S r
        bq = dot(vtmp.p, kdv, n);    // (S r) . K (S r)
    };
    auto objective = [&](const double *sv, double &ah, double &bh, double &al, double &bl) {
        rho_terms(sv, rhi.p, kd.p, ah, bh);
        rho_terms(sv, rlo, kd2.p, al, bl);
        return (ah > 0 && bh > 0 && al > 0 && bl > 0) ? std::log(bh / ah) - std::log(bl / al) : 0.0;
    };
    // r_hi: power iterations of K S from the right-hand side
    mlp_infer(rhs, yprev.p);
    HIPCHK(hipMemcpyAsync(rhi.p, rhs, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    for (int it = 0; it < 6; it++) {
        kprod(lminv.p, rhi.p, kd.p);
        const double nrm = std::sqrt(dot(kd.p, kd.p, n));
        if (!(nrm > 0.0)) break;
        v2g::vec_axpby(n, 1.0 / nrm, kd.p, 0.0, nullptr, rhi.p, st);
    }
    // the base diagonal (g = 0) on the same probes, for the reported change
    if (!mlp_jacobi_base && !ones.p) {
        ones.alloc(n);
        v2g::vec_fill(n, 1.0, ones.p, st);
    }
    double ah, bh, al, bl;
    const double l0 = objective(mlp_jacobi_base ? jacobi.p : ones.p, ah, bh, al, bl);
    std::uniform_int_distribution<int64_t> ud(0, n - 1);
    std::vector<int64_t> idx(nb);
    double l = l0;
    for (int k = 0; k < steps; k++) {
        mlp_infer(rhs, yprev.p);              // s = lminv, scales[0..1]
        l = objective(lminv.p, ah, bh, al, bl);
        double ab[4] = {ah, bh, al, bl};
        HIPCHK(hipMemcpyAsync(scales.p + 2, ab, sizeof(ab), hipMemcpyHostToDevice, st));
        for (auto &i : idx) i = ud(rng);
        HIPCHK(hipMemcpyAsync(bidx.p, idx.data(), sizeof(int64_t) * nb, hipMemcpyHostToDevice, st));
        v2g::mlp_batch_spec(nb, bidx.p, code.p, rownorm.p, rhs, yprev.p, rhi.p, kd.p, rlo, kd2.p,
                            lminv.p, scales.p, feat.p, wt.p, 3.0, (double)n / nb, mlp_jacobi_base, pen,
                            st);
        v2g::mlp_grad(nb, feat.p, nullptr, wt.p, W.p, mlp_part.p, mlp_lpart.p, grad.p, dloss.p, st);
        v2g::mlp_adam(W.p, grad.p, am.p, av.p, lr, ++adam_t, st);
    }
    HIPCHK(hipStreamSynchronize(st));
    return l - l0;
}

double v2g_solver::dot(const double *a, const double *b, int64_t n)
{
    v2g::vec_dot_partial(n, a, b, red.p, st);
    v2g::vec_reduce(red.p, red.p + v2g::kRedBlocks, st);
    double out = 0.0;
    HIPCHK(hipMemcpyAsync(&out, red.p + v2g::kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return out;
}

int v2g_solver::cg(const double *rhs, double *y, double tolv, int maxiter, int precond,
                   int flexible, double *res)
{
    const double *minv = nullptr;
    if (precond == 1) { ensure_jacobi(); minv = jacobi.p; }
    else if (precond == 2) {
        if (!have_user) throw std::invalid_argument("precond 2 needs v2g_set_precond_diag");
        minv = user_minv.p;
    } else if (precond == 4) {
        minv = lminv.p;  // the learned diagonal from mlp_infer
    }
    return cg_with(own_cg_bufs(), use_graphs, rhs, y, tolv, maxiter, minv, flexible, res);
}

int v2g_solver::cg_with(const CgBufs &B, bool graphs, const double *rhs, double *y, double tolv,
                        int maxiter, const double *minv, int flexible, double *res)
{
    const int64_t n = L->ncon;
    // Ap = A A^T y ; r = b - Ap ; p = M^-1 r   (cg_solver.cc:141-150)
    v2g::launch_atu(plan, y, B.t, st);
    v2g::launch_au(plan, B.t, B.Ap, st);
    v2g::cg_init(n, rhs, B.Ap, B.r, B.p, minv, nullptr, B.partial, B.scal, tolv, maxiter, st);
    // one chunk of kCgChunk (even) iterations; every kernel argument is fixed,
    // the iteration number and the done flags live in scal
    auto enqueue_chunk = [&]() {
        const bool fused = fused_cg();
        for (int c = 0; c < kCgChunk; c++) {
            const int par = c & 1;
            const double *done = B.scal + v2g::CG_DONE0 + par;
            if (fused) {
                // t = A^T p with the tile squares -> alpha; r -= alpha A t in
                // the A.x launch (A p is never stored) with the r.z, r.r
                // partials; then y, the test and p as below
                double *stepp = B.rowpart + plan.x_tiles;
                v2g::launch_atu(plan, B.p, B.t, st, done, B.rowpart);
                v2g::cg_alpha_from_tiles(plan.x_tiles, B.rowpart, B.scal, par, st);
                v2g::CgFuse fz;
                fz.r = B.r;
                fz.minv = minv;
                fz.scal = B.scal;
                fz.part = stepp;
                fz.ntiles = plan.row_tiles;
                v2g::launch_au(plan, B.t, nullptr, st, done, nullptr, nullptr, fz);
                v2g::cg_step_from_tiles(plan.row_tiles, stepp, B.partial, B.scal, par, st);
            } else {
                v2g::launch_atu(plan, B.p, B.t, st, done);
                v2g::launch_au(plan, B.t, B.Ap, st, done, B.p, B.rowpart);   // Ap and p.Ap
                v2g::cg_dot_from_rows(plan.row_tiles, B.rowpart, B.partial, B.scal, par, st);
                v2g::cg_step(n, B.r, B.Ap, minv, B.partial, B.scal, par, st);
            }
            v2g::cg_finish_iter(n, y, B.r, B.p, minv, B.partial, B.scal, par, flexible, st);
        }
    };
    hipGraphExec_t exec = nullptr;
    if (graphs && n <= kGraphMaxRows) {
        if (!(cg_exec && cg_exec_y == y && cg_exec_minv == minv && cg_exec_flex == flexible)) {
            if (cg_exec) { (void)hipGraphExecDestroy(cg_exec); cg_exec = nullptr; }
            // relaxed mode: the other host threads of a bond scan keep allocating,
            // copying and synchronizing their own streams while this one captures
            hipGraph_t g = nullptr;
            bool ok = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) == hipSuccess;
            if (ok) {
                enqueue_chunk();
                ok = hipStreamEndCapture(st, &g) == hipSuccess && g;
            }
            if (ok) ok = hipGraphInstantiate(&cg_exec, g, nullptr, nullptr, 0) == hipSuccess;
            if (g) (void)hipGraphDestroy(g);
            if (!ok) {            // capture unsupported here: stay with plain launches
                cg_exec = nullptr;
                use_graphs = false;
                (void)hipGetLastError();
            } else {
                cg_exec_y = y;
                cg_exec_minv = minv;
                cg_exec_flex = flexible;
            }
        }
        exec = cg_exec;
    }
    double hs[v2g::CG_NSLOTS];
    int it = 0;
    for (;;) {
        if (exec) HIPCHK(hipGraphLaunch(exec, st));
        else enqueue_chunk();
        it += kCgChunk;
        HIPCHK(hipMemcpyAsync(hs, B.scal, sizeof(hs), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (hs[v2g::CG_DONE0 + (it & 1)] != 0.0) break;
        if (it > maxiter + 2 * kCgChunk) break;  // defensive: cannot happen
    }
    if (res) *res = hs[v2g::CG_NRM];
    last_cg_nrm0 = hs[v2g::CG_NRM0];
    return (int)hs[v2g::CG_ITER];
}

// warm_sweep_cap > 0 bounds the sweeps of a warm-started call: inside the BPSDP
// loop the eigenvectors keep improving from call to call, so the projection
// becomes exact at the fixed point while early calls stay cheap.  x is always
// rebuilt PSD from the current V; z = mu x - M keeps M exact.
void v2g_solver::update_xz(double mu, const double *y, const double *c, double *x, double *zv,
                           int warm_sweep_cap, int inner_sweep_cap, int cross_order)
{
    v2g::launch_atu(plan, y, ATy.p, st);
    v2g::eig_small(d_small.p, (int)small.size(), ATy.p, c, x, zv, mu, st);
    last_sweeps = 0;
    if (large.empty()) return;
    const int cold = warm ? 0 : 1;
    // autotuning: this warm call runs candidate tune_k / kTuneReps
    const bool timed = tuning && !cold;
    if (timed) {
        const auto &cd = tune_cand[tune_k / kTuneReps];
        set_shape(cd.first, cd.second);
    }
    const int nl = (int)large.size();
    v2g::eig_build_m(d_large.p, nl, max_N, ATy.p, c, x, mu, Mw.p, Aw.p, Vw.p, cold, st);
    // a split that may stop after capped sweeps (in-loop tracking) rotates
    // above 2e-14 ||M||_F; a split that must converge uses the tight threshold
    const bool capped = !cold && (warm_sweep_cap > 0 || inner_sweep_cap > 0);
    v2g::eig_tol(d_large.p, nl, Mw.p, tol.p, frop.p, capped ? 0 : 1, st);
    if (!cold) {
        // A = V^T (M V)
        if (!d_gemm_warm1.p) {
            HIPCHK(hipStreamSynchronize(st));
            std::vector<v2g::GemmDesc> g1, g2;
            for (const auto &b : large) {
                g1.push_back({Mw.p + b.woff, Vw.p + b.woff, Tw.p + b.woff, b.N, b.N, b.N, b.N, b.N,
                              b.N, 0, 0, 0});
                g2.push_back({Vw.p + b.woff, Tw.p + b.woff, Aw.p + b.woff, b.N, b.N, b.N, b.N, b.N,
                              b.N, 1, 0, 1});
            }
            d_gemm_warm1.upload(g1);
            d_gemm_warm2.upload(g2);
        }
        v2g::gemm_batched(d_gemm_warm1.p, nl, max_N, max_N, st);
        v2g::gemm_batched(d_gemm_warm2.p, nl, max_N, max_N, st);
    }
    const int maxsweeps = (!cold && warm_sweep_cap > 0) ? warm_sweep_cap : 40;
    const int inner = (!cold && inner_sweep_cap > 0) ? inner_sweep_cap : v2g::kEigFullInner;
    int sweeps = 0, swept = 0;
    if (timed) HIPCHK(hipEventRecord(ev_t0, st));
    for (; sweeps < maxsweeps; sweeps++) {
        swept++;
        HIPCHK(hipMemsetAsync(counter.p, 0, sizeof(int), st));
        if (nlanes > 2) {   // lanes 2, 4, ... start from the engine stream's state
            HIPCHK(hipEventRecord(ev_fork, st));
            for (int l = 2; l < nlanes; l += 2) HIPCHK(hipStreamWaitEvent(lane_st[l], ev_fork, 0));
        }
        for (int rnd = 0; rnd < max_nb - 1; rnd++)
            for (int l = 0; l < nlanes; l++) {
                if (rnd >= lane_nb[l] - 1) continue;
                const v2g::EigBlock *bl = d_large.p + lane_off[l];
                hipStream_t s = l ? lane_st[l] : st;
                // after its first round a sweep takes the cross-pair order
                // (every pair is still visited once per sweep): always in a
                // split that must converge, by option in a capped tracking sweep
                const int bip = ((capped ? cross_order : 1) && rnd > 0) ? 1 : 0;
                v2g::eig_pairs(bl, lane_n[l], lane_pairs[l], rnd, Aw.p, Qbuf.p, active.p,
                               tol.p + lane_off[l], counter.p, inner, pair_wide ? 1 : 0, bip, s);
                if (rnd == 0 && l % 2 == 0 && l + 1 < nlanes) {
                    // lane l + 1 starts once lane l's first pair problems are
                    // done, so in steady state each of the two runs its tile
                    // updates while the other solves its pair problems
                    HIPCHK(hipEventRecord(ev_start[l], s));
                    HIPCHK(hipStreamWaitEvent(lane_st[l + 1], ev_start[l], 0));
                }
                v2g::eig_apply(bl, lane_n[l], lane_tiles[l], rnd, Aw.p, Vw.p, Qbuf.p, active.p, s);
            }
        for (int l = 1; l < nlanes; l++) {
            HIPCHK(hipEventRecord(ev_join[l], lane_st[l]));
            HIPCHK(hipStreamWaitEvent(st, ev_join[l], 0));
        }
        if (timed) HIPCHK(hipEventRecord(ev_t1, st));
        int cnt = 0;
        HIPCHK(hipMemcpyAsync(&cnt, counter.p, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (cnt == 0) break;
    }
    if (timed) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, ev_t0, ev_t1));
        ms /= (float)std::max(1, swept);
        float &best = tune_ms[tune_k / kTuneReps];
        if (tune_k % kTuneReps == 0 || ms < best) best = ms;
        if (++tune_k == (int)tune_cand.size() * kTuneReps) {
            int bi = 0;
            for (int i = 1; i < (int)tune_cand.size(); i++)
                if (tune_ms[i] < tune_ms[bi]) bi = i;
            tuning = false;
            set_shape(tune_cand[bi].first, tune_cand[bi].second);
        }
    }
    last_sweeps = sweeps;
    last_split_capped = !cold && sweeps > 0 && (sweeps >= maxsweeps || inner < v2g::kEigFullInner);
    // x = (V s)(V s)^T with s = sqrt(max(lambda,0)/mu); z = mu x - M
    v2g::eig_scale(d_large.p, nl, Aw.p, Vw.p, Tw.p, mu, st);
    // the reconstruction descriptors point into x: built once per x buffer
    // (an upload frees and allocates, which would drain the device every call)
    int maxn = 0;
    for (const auto &b : large) maxn = std::max(maxn, b.n);
    if (rec_x != x || !d_gemm_rec.p) {
        std::vector<v2g::GemmDesc> g3;
        for (const auto &b : large)
            g3.push_back({Tw.p + b.woff, Tw.p + b.woff, x + b.xoff, b.n, b.n, b.N, b.N, b.N, b.n, 0,
                          1, 1});
        HIPCHK(hipStreamSynchronize(st));
        d_gemm_rec.upload(g3);
        rec_x = x;
    }
    v2g::gemm_batched(d_gemm_rec.p, nl, maxn, maxn, st);
    v2g::eig_z(d_large.p, nl, x, Mw.p, zv, mu, st);
    warm = true;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------


namespace {

// host copy of the primal block `b` of irrep `hh` (n x n entries)
std::vector<double> block_to_host(v2g_handle h, const double *x, int b, int hh, int64_t nn)
{
    std::vector<double> v((size_t)nn);
    const int64_t o = h->L->xoff[b][hh];
    if (nn == 0) return v;
    if (o < 0) throw std::invalid_argument("primal block absent from this layout");
    HIPCHK(hipMemcpy(v.data(), x + o, sizeof(double) * nn, hipMemcpyDeviceToHost));
    return v;
}

}  // namespace

extern "C" {

int v2g_create(const v2g_layout_desc *desc, v2g_handle *out)
{
    API_BEGIN
    if (!desc || !out) throw std::invalid_argument("null argument");
    *out = new v2g_solver(*desc);
    API_END
}

int v2g_destroy(v2g_handle h)
{
    API_BEGIN
    delete h;
    API_END
}

int v2g_dims(v2g_handle h, long long *dimx, long long *ncon, int *nblocks)
{
    API_BEGIN
    if (!h) throw std::invalid_argument("null handle");
    if (dimx) *dimx = h->L->dimx;
    if (ncon) *ncon = h->L->ncon;
    if (nblocks) *nblocks = (int)h->L->block_dims.size();
    API_END
}

int v2g_block_dims(v2g_handle h, int *dims)
{
    API_BEGIN
    if (!h || !dims) throw std::invalid_argument("null argument");
    std::copy(h->L->block_dims.begin(), h->L->block_dims.end(), dims);
    API_END
}

int v2g_set_stream(v2g_handle h, void *s)
{
    API_BEGIN
    if (!h) throw std::invalid_argument("null handle");
    h->st = s ? (hipStream_t)s : h->own;
    API_END
}

int v2g_Au(v2g_handle h, const double *x, double *Ax)
{
    API_BEGIN
    if (!h || !x || !Ax) throw std::invalid_argument("null argument");
    v2g::launch_au(h->plan, x, Ax, h->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));
    API_END
}

int v2g_ATu(v2g_handle h, const double *y, double *ATy)
{
    API_BEGIN
    if (!h || !y || !ATy) throw std::invalid_argument("null argument");
    v2g::launch_atu(h->plan, y, ATy, h->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));
    API_END
}

int v2g_cg_Ax(v2g_handle h, const double *u, double *out)
{
    API_BEGIN
    if (!h || !u || !out) throw std::invalid_argument("null argument");
    v2g::launch_atu(h->plan, u, h->t.p, h->st);
    v2g::launch_au(h->plan, h->t.p, out, h->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));
    API_END
}

int v2g_row_norm2(v2g_handle h, double *d)
{
    API_BEGIN
    if (!h || !d) throw std::invalid_argument("null argument");
    v2g::launch_row_norm2(h->plan, d, h->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));
    API_END
}

int v2g_cg_solve(v2g_handle h, const double *rhs, double *y, double tol, int maxiter, int precond,
                 int flexible, int *iters, double *resnorm)
{
    API_BEGIN
    if (!h || !rhs || !y) throw std::invalid_argument("null argument");
    if (precond < 0 || precond > 2) throw std::invalid_argument("precond must be 0, 1 or 2");
    const int it = h->cg(rhs, y, tol, maxiter, precond, flexible, resnorm);
    if (iters) *iters = it;
    HIPCHK(hipGetLastError());
    API_END
}

int v2g_cg_solve_noalloc(v2g_handle h, const double *rhs, double *y, double tol, int maxiter,
                         int precond, int flexible, int *iters, double *resnorm)
{
    API_BEGIN
    if (!h || !rhs || !y) throw std::invalid_argument("null argument");
    if (precond != 0 && precond != 2)
        throw std::invalid_argument("the allocation-free CG takes precond 0 or 2");
    if (precond == 2 && !h->have_user) throw std::invalid_argument("precond 2 needs v2g_set_precond_diag");
    // scratch layout: r, p, Ap (ncon each), t (dimx), partial sums, scalars,
    // per-tile partials of the CG reductions; every region starts on a 256-byte
    // boundary
    auto up = [](int64_t n) { return (n + 31) / 32 * 32; };
    const int64_t ncon = h->L->ncon, dimx = h->L->dimx;
    const int64_t need = 3 * up(ncon) + up(dimx) + up(v2g::kPartialArrays * v2g::kRedBlocks) +
                         up(v2g::CG_NSLOTS) + up(h->rowpart_len());
    if (need > (int64_t)kScratchDoubles)
        throw std::invalid_argument("problem too large for the allocation-free CG: needs " +
                                    std::to_string(need) + " doubles of scratch, at most " +
                                    std::to_string(kScratchDoubles) + " (64 MiB) are available");
    double *s = h->scratch_dev;
    v2g_solver::CgBufs B{};
    B.r = s;
    B.p = B.r + up(ncon);
    B.Ap = B.p + up(ncon);
    B.t = B.Ap + up(ncon);
    B.partial = B.t + up(dimx);
    B.scal = B.partial + up(v2g::kPartialArrays * v2g::kRedBlocks);
    B.rowpart = B.scal + up(v2g::CG_NSLOTS);
    const double *minv = precond == 2 ? h->user_minv.p : nullptr;
    const int it = h->cg_with(B, false, rhs, y, tol, maxiter, minv, flexible, resnorm);
    if (iters) *iters = it;
    HIPCHK(hipGetLastError());
    API_END
}

int v2g_set_precond_diag(v2g_handle h, const double *minv)
{
    API_BEGIN
    if (!h || !minv) throw std::invalid_argument("null argument");
    if (!h->user_minv.p) h->user_minv.alloc(h->L->ncon);
    HIPCHK(hipMemcpyAsync(h->user_minv.p, minv, sizeof(double) * h->L->ncon,
                          hipMemcpyDeviceToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    h->have_user = true;
    API_END
}

int v2g_update_xz(v2g_handle h, double mu, const double *y, const double *c, double *x, double *z)
{
    API_BEGIN
    if (!h || !y || !c || !x || !z) throw std::invalid_argument("null argument");
    if (!(mu > 0.0)) throw std::invalid_argument("mu must be positive");
    h->update_xz(mu, y, c, x, z);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));
    API_END
}

int v2g_reset_eig(v2g_handle h)
{
    API_BEGIN
    if (!h) throw std::invalid_argument("null handle");
    h->warm = false;
    API_END
}

int v2g_set_eig_shape(v2g_handle h, int lanes, int pair_threads)
{
    API_BEGIN
    if (!h) throw std::invalid_argument("null handle");
    if (lanes != 0 && lanes != 1 && lanes != 2) throw std::invalid_argument("lanes must be 0, 1 or 2");
    if (pair_threads != 0 && pair_threads != 256 && pair_threads != 1024)
        throw std::invalid_argument("pair_threads must be 0, 256 or 1024");
    h->tuning = false;
    h->set_shape(lanes, pair_threads == 0 ? -1 : (pair_threads == 1024 ? 1 : 0));
    API_END
}

int v2g_get_eig_shape(v2g_handle h, int *lanes, int *pair_threads)
{
    API_BEGIN
    if (!h) throw std::invalid_argument("null handle");
    if (lanes) *lanes = h->nlanes;
    if (pair_threads) *pair_threads = h->pair_wide ? 1024 : 256;
    API_END
}

int v2g_set_autotune(v2g_handle h, int on)
{
    API_BEGIN
    if (!h) throw std::invalid_argument("null handle");
    if (on) h->start_tuning();
    else h->tuning = false;
    API_END
}

int v2g_set_iter_callback(v2g_handle h, v2g_iter_callback cb, void *user)
{
    API_BEGIN
    if (!h) throw std::invalid_argument("null handle");
    h->cb = cb;
    h->cb_user = user;
    API_END
}

int v2g_bpsdp(v2g_handle h, const v2g_params *prm, const double *c, const double *b, double *x,
              double *z, double *y, v2g_stats *out)
{
    API_BEGIN
    if (!h || !prm || !c || !b || !x || !z || !y) throw std::invalid_argument("null argument");
    const int64_t ncon = h->L->ncon, dimx = h->L->dimx;
    hipStream_t st = h->st;
    const double tau = prm->tau;
    // a resumed solve (oiter0 > 0) continues with the mu, errors and iteration
    // count of the call it follows
    double mu = prm->mu0 > 0.0 ? prm->mu0 : 1.0, ep = prm->ep0, ed = prm->ed0, egap = 0.0,
           ecur = 0.0, edual = 0.0;
    if (prm->oiter0 < 0) throw std::invalid_argument("oiter0 must be >= 0");
    int oiter = prm->oiter0;
    const int oiter_first = oiter;
    int exact_splits = 0;
    // the in-loop split tracks the eigenvectors with capped sweeps; before the
    // stopping rule may accept an iteration, that iteration's split is redone
    // to convergence from the same x, so the answer is the reference's fixed point
    const bool capped_loop = prm->eig_warm_sweeps > 0 || prm->eig_inner_sweeps > 0;
    if (capped_loop && !h->xprev.p) h->xprev.alloc(dimx);
    long long cgtot = 0;
    double tcg = 0.0, txz = 0.0;
    const double t0 = now();
    v2g_stats s{};
    bool stop = false;
    const bool learned = prm->precond == 3;
    if (learned) {
        h->ensure_mlp();
        h->mlp_jacobi_base = prm->mlp_base == 1 ? 1 : 0;
    }
    // the learned diagonal's base method: plain CG (identity) or Jacobi-CG
    const int base_pc = (learned && h->mlp_jacobi_base) ? 1 : 0;
    std::mt19937_64 rng(20240607ULL);
    long long cg_train = 0;
    int n_train = 0;
    // the A/B guard state lives in the handle, so a later call keeps its decision
    bool &fallback = h->mlp_fallback, &first_learned = h->mlp_first_learned;
    int &last_base_iters = h->mlp_last_base_iters;
    double last_loss = 0.0;
    do {
        const double t1 = now();
        // Ax = -tau mu (A x - b); z = c - z; B = A z + Ax  (v2rdm_solver.cc:1784-1794).
        // After the first iteration of a call, Ax already holds A x - b for this x
        // (the error evaluation of the previous iteration)
        if (oiter == oiter_first) {
            v2g::launch_au(h->plan, x, h->Ax.p, st);
            v2g::vec_axpby(ncon, -tau * mu, h->Ax.p, tau * mu, b, h->Ax.p, st);
        } else {
            v2g::vec_axpby(ncon, -tau * mu, h->Ax.p, 0.0, h->Ax.p, h->Ax.p, st);
        }
        v2g::vec_axpby(dimx, 1.0, c, -1.0, z, z, st);
        v2g::launch_au(h->plan, z, h->Bv.p, st);
        v2g::vec_axpby(ncon, 1.0, h->Bv.p, 1.0, h->Ax.p, h->Bv.p, st);
        double cgtol;  // v2rdm_solver.cc:1797-1805
        if (oiter == 0 || (oiter == oiter_first && !(ep > 0.0 && ed > 0.0))) cgtol = 0.01;
        else cgtol = (ep > ed) ? 0.01 * ed : 0.01 * ep;
        if (cgtol < prm->cg_convergence) cgtol = prm->cg_convergence;
        double res = 0.0;
        int pc = prm->precond, flex = prm->flexible;
        // training spans the first train_iters outer iterations of the
        // handle's life (it carries over between calls, like the weights)
        const bool training = learned && h->mlp_outer_trained < prm->train_iters;
        if (learned) {
            if (training || fallback) {
                pc = base_pc;
                flex = 0;
            } else {
                h->mlp_infer(h->Bv.p, y);
                pc = 4;
                flex = 1;
            }
        }
        // CG starting guess: the reference restarts from y_k; with
        // cg_extrapolate the guess is 2 y_k - y_(k-1), which cuts the CG work
        // without changing the system, the tolerance or the outer iteration
        if (prm->cg_extrapolate) {
            if (!h->yhist.p) h->yhist.alloc(ncon);
            if (oiter == oiter_first)
                HIPCHK(hipMemcpyAsync(h->yhist.p, y, sizeof(double) * ncon, hipMemcpyDeviceToDevice,
                                      st));
            else
                v2g::vec_extrapolate(ncon, y, h->yhist.p, st);
        }
        if (training)
            HIPCHK(hipMemcpyAsync(h->yprev.p, y, sizeof(double) * ncon, hipMemcpyDeviceToDevice, st));
        const int iiter = h->cg(h->Bv.p, y, cgtol, prm->cg_maxiter, pc, flex, &res);
        cgtot += iiter;
        // convergence rate of this solve: ln(||r_0|| / ||r||) per iteration
        const double rate = (iiter > 0 && res > 0.0 && h->last_cg_nrm0 > res)
                                ? std::log(h->last_cg_nrm0 / res) / iiter : -1.0;
        if (training) {
            // the base solve's final residual carries its slowest modes
            last_loss = h->mlp_train(h->Bv.p, h->r.p, prm->train_batch > 0 ? prm->train_batch : 16384,
                                     prm->train_steps > 0 ? prm->train_steps : 10,
                                     prm->train_lr > 0 ? prm->train_lr : 2e-3, prm->train_penalty,
                                     rng);
            cg_train += iiter;
            n_train++;
            h->mlp_outer_trained++;
            last_base_iters = iiter;
            // the base method's rate over the second half of the training phase
            if (rate > 0.0 && 2 * h->mlp_outer_trained > prm->train_iters) {
                h->base_rate_sum += rate;
                h->base_rate_n++;
            }
        } else if (pc == 4) {
            // the true residual decides: abandon the learned diagonal if a
            // solve fails to converge, or if over its first kGuard solves it
            // converges clearly more slowly per iteration than its base did
            constexpr int kGuard = 4;
            if (iiter >= prm->cg_maxiter || !std::isfinite(res)) fallback = true;
            // a solve twice as long as the base method's last one ends it at once
            if (first_learned && iiter > 2 * last_base_iters + 8) fallback = true;
            if (first_learned && rate > 0.0) {
                h->learned_rate_sum += rate;
                if (++h->learned_rate_n == kGuard) {
                    first_learned = false;
                    if (h->base_rate_n > 0 && h->learned_rate_sum / kGuard <
                                                  0.9 * h->base_rate_sum / h->base_rate_n)
                        fallback = true;
                }
            }
        }
        s.precond_active = pc == 4 ? 3 : pc;
        const double t2 = now();
        tcg += t2 - t1;
        if (capped_loop)
            HIPCHK(hipMemcpyAsync(h->xprev.p, x, sizeof(double) * dimx, hipMemcpyDeviceToDevice, st));
        h->update_xz(mu, y, c, x, z, prm->eig_warm_sweeps, prm->eig_inner_sweeps,
                     prm->eig_pair_order);
        HIPCHK(hipStreamSynchronize(st));
        txz += now() - t2;
        // ed = ||A^T y - c + z||, ep = ||A x - b||  (v2rdm_solver.cc:1829-1837).
        // Every call follows an update_xz with this y, which left A^T y in ATy
        // The four sums are formed on the device and read back together.
        auto errors = [&]() {
            constexpr int64_t R = v2g::kRedBlocks;
            double *red = h->red.p;
            v2g::vec_lincomb_sq(dimx, 1.0, h->ATy.p, 1.0, z, -1.0, c, h->ATy.p, red, st);
            v2g::launch_au(h->plan, x, h->Ax.p, st);
            v2g::vec_lincomb_sq(ncon, 1.0, h->Ax.p, -1.0, b, 0.0, nullptr, h->Ax.p, red + R, st);
            v2g::vec_dot_partial(dimx, c, x, red + 2 * R, st);
            v2g::vec_dot_partial(ncon, b, y, red + 3 * R, st);
            for (int q = 0; q < 4; q++) v2g::vec_reduce(red + q * R, red + 4 * R + q, st);
            double o[4];
            HIPCHK(hipMemcpyAsync(o, red + 4 * R, sizeof(o), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            ed = std::sqrt(o[0]);
            ep = std::sqrt(o[1]);
            ecur = o[2];
            edual = o[3];
        };
        errors();
        auto passes = [&]() {
            return !(ep > prm->r_convergence || ed > prm->r_convergence ||
                     std::fabs(ecur - edual) > prm->e_convergence);
        };
        if (passes() && h->last_split_capped) {
            // the tracked split rotated in its last sweep: redo this
            // iteration's split to convergence (same y, same x on entry)
            const double t3 = now();
            HIPCHK(hipMemcpyAsync(x, h->xprev.p, sizeof(double) * dimx, hipMemcpyDeviceToDevice, st));
            h->update_xz(mu, y, c, x, z, 0, 0);
            HIPCHK(hipStreamSynchronize(st));
            txz += now() - t3;
            exact_splits++;
            errors();
        }
        if (oiter % prm->mu_update_frequency == 0 && oiter > 0) mu = mu * ep / ed;
        if (prm->verbose)
            std::fprintf(stderr, "      %5i %5i %14.8f %14.8f %11.4e %7.3f %10.5e %10.5e  sweeps %d\n",
                         oiter, iiter, ecur, edual, std::fabs(ecur - edual), mu, ep, ed,
                         h->last_sweeps);
        oiter++;
        egap = std::fabs(ecur - edual);
        s.oiter = oiter;
        s.cg_iters = cgtot;
        s.energy_primal = ecur;
        s.energy_dual = edual;
        s.ep = ep;
        s.ed = ed;
        s.mu = mu;
        s.t_cg = tcg;
        s.t_xz = txz;
        s.t_total = now() - t0;
        s.eig_sweeps = h->last_sweeps;
        s.cg_iters_train = cg_train;
        s.train_outer = n_train;
        s.mlp_loss = last_loss;
        s.fallback = fallback ? 1 : 0;
        s.exact_splits = exact_splits;
        if (h->cb && h->cb(h->cb_user, oiter, &s)) stop = true;
        if (oiter >= prm->maxiter) break;
        if (prm->max_seconds > 0 && now() - t0 > prm->max_seconds) break;
    } while (!stop && (ep > prm->r_convergence || ed > prm->r_convergence || egap > prm->e_convergence));
    s.converged = !(ep > prm->r_convergence || ed > prm->r_convergence || egap > prm->e_convergence);
    s.t_total = now() - t0;
    if (out) *out = s;
    HIPCHK(hipGetLastError());
    API_END
}

int v2g_build_b(v2g_handle h, double *b)
{
    API_BEGIN
    if (!h || !b) throw std::invalid_argument("null argument");
    const v2g::Layout &L = *h->L;
    std::fill(b, b + L.ncon, 0.0);
    const double na = L.na, nb = L.nb, n = na + nb;
    if (L.opt.sz) {
        b[L.rowoff[v2g::RF_TRACE][0][0]] = na * nb;
        b[L.rowoff[v2g::RF_TRACE][1][0]] = na * (na - 1.0);
        b[L.rowoff[v2g::RF_TRACE][2][0]] = nb * (nb - 1.0);
    } else {
        b[L.rowoff[v2g::RF_TRACE][3][0]] = n * (n - 1.0);
    }
    for (int s = 0; s < 2; s++)
        for (int hh = 0; hh < L.nirrep; hh++) {
            const int m = L.opt.amopi[hh];
            const int64_t o = L.rowoff[v2g::RF_D1Q1][s][hh];
            for (int i = 0; i < m; i++) b[o + (int64_t)i * m + i] = 1.0;
        }
    if (L.opt.spin) {
        const double ms = (L.opt.multiplicity - 1.0) / 2.0;
        b[L.rowoff[v2g::RF_S2][0][0]] = 0.5 * (na + nb) + ms * ms - ms * (ms + 1.0);
    }
    if (L.q2_plain()) {  // spin-adapted Q2 rows have b = 0 (v2rdm_solver.cc:2709-2745)
        for (int hh = 0; hh < L.nirrep; hh++) {
            const int64_t g = L.gems_ab[hh], o = L.rowoff[v2g::RF_Q2AB][0][hh];
            for (int64_t ij = 0; ij < g; ij++) b[o + ij * g + ij] = -1.0;
        }
        for (int s = 0; s < 2; s++)
            for (int hh = 0; hh < L.nirrep; hh++) {
                const int64_t g = L.gems_aa[hh], o = L.rowoff[v2g::RF_Q2AA][s][hh];
                for (int64_t ij = 0; ij < g; ij++) b[o + ij * g + ij] = -1.0;
            }
    }
    API_END
}

int v2g_build_c(v2g_handle h, const double *h1, const double *eri, double *c)
{
    API_BEGIN
    if (!h || !h1 || !eri || !c) throw std::invalid_argument("null argument");
    const v2g::Layout &L = *h->L;
    const int64_t a = L.amo;
    std::fill(c, c + L.dimx, 0.0);
    auto E = [&](int p, int q, int r, int s) { return eri[((p * a + q) * a + r) * a + s]; };
    for (int hh = 0; hh < L.nirrep; hh++) {
        const int64_t g = L.gems_ab[hh];
        const int *bp = L.bab.data() + L.bab_off[hh];
        for (int64_t ij = 0; ij < g; ij++) {
            const int i = bp[ij] & 0xffff, j = bp[ij] >> 16;
            for (int64_t kl = 0; kl < g; kl++) {
                const int k = bp[kl] & 0xffff, l = bp[kl] >> 16;
                c[L.xoff[v2g::XB_D2AB][hh] + ij * g + kl] = E(i, k, j, l);
            }
        }
        const int64_t ga = L.gems_aa[hh];
        const int *ap = L.baa.data() + L.baa_off[hh];
        for (int64_t ij = 0; ij < ga; ij++) {
            const int i = ap[ij] & 0xffff, j = ap[ij] >> 16;
            for (int64_t kl = 0; kl < ga; kl++) {
                const int k = ap[kl] & 0xffff, l = ap[kl] >> 16;
                const double v = E(i, k, j, l) - E(i, l, j, k);
                c[L.xoff[v2g::XB_D2AA][hh] + ij * ga + kl] = v;
                c[L.xoff[v2g::XB_D2BB][hh] + ij * ga + kl] = v;
            }
        }
        const int m = L.opt.amopi[hh];
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) {
                const double v = h1[(int64_t)(i + L.poff[hh]) * a + (j + L.poff[hh])];
                c[L.xoff[v2g::XB_D1A][hh] + (int64_t)i * m + j] = v;
                c[L.xoff[v2g::XB_D1B][hh] + (int64_t)i * m + j] = v;
            }
    }
    API_END
}

int v2g_build_b_device(v2g_handle h, double *b)
{
    API_BEGIN
    if (!h || !b) throw std::invalid_argument("null argument");
    const v2g::Layout &L = *h->L;
    v2g::build_b_device(h->d_lay.p, h->d_rowsegs.p, (int)L.rowsegs.size(), L.ncon, b, h->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));
    API_END
}

int v2g_build_c_device(v2g_handle h, const double *h1, const double *eri, double *c)
{
    API_BEGIN
    if (!h || !h1 || !eri || !c) throw std::invalid_argument("null argument");
    const v2g::Layout &L = *h->L;
    int mgab = 0, mgaa = 0, mm = 0;
    for (int hh = 0; hh < L.nirrep; hh++) {
        mgab = std::max(mgab, L.gems_ab[hh]);
        mgaa = std::max(mgaa, L.gems_aa[hh]);
        mm = std::max(mm, L.opt.amopi[hh]);
    }
    v2g::build_c_device(h->d_lay.p, L.nirrep, mgab, mgaa, mm, L.dimx, h1, eri, c, h->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));
    API_END
}

int v2g_guess_random_device(v2g_handle h, double *x, double *z, double *y)
{
    API_BEGIN
    if (!h || !x || !z || !y) throw std::invalid_argument("null argument");
    const v2g::Layout &L = *h->L;
    const int64_t total = 2 * L.dimx + L.ncon;
    const std::vector<uint32_t> starts = v2g::guess_chunk_starts(total);
    DevBuf<uint32_t> d_starts;
    d_starts.upload(starts);
    v2g::guess_device(d_starts.p, total, L.dimx, x, z, y, h->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));
    API_END
}

int v2g_guess_random(v2g_handle h, double *x, double *z, double *y)
{
    API_BEGIN
    if (!h || !x || !z || !y) throw std::invalid_argument("null argument");
    const v2g::Layout &L = *h->L;
    srand(0);
    for (int64_t i = 0; i < L.dimx; i++) {
        x[i] = ((double)rand() / RAND_MAX - 1.0) * 2.0;
        z[i] = ((double)rand() / RAND_MAX - 1.0) * 2.0;
    }
    for (int64_t i = 0; i < L.ncon; i++) y[i] = ((double)rand() / RAND_MAX - 1.0) * 2.0;
    API_END
}

// RDM outputs (host side; the RDMs are small next to x)

int v2g_get_opdm(v2g_handle h, const double *x, double *opdm)
{
    API_BEGIN
    if (!h || !x || !opdm) throw std::invalid_argument("null argument");
    HIPCHK(hipStreamSynchronize(h->st));
    const v2g::Layout &L = *h->L;
    int64_t o = 0;
    for (int hh = 0; hh < L.nirrep; hh++) {
        const int64_t n = L.opt.amopi[hh], nn = n * n;
        const auto a = block_to_host(h, x, v2g::XB_D1A, hh, nn);
        const auto b = block_to_host(h, x, v2g::XB_D1B, hh, nn);
        for (int64_t i = 0; i < nn; i++) opdm[o + i] = a[i] + b[i];
        o += nn;
    }
    API_END
}

int v2g_get_tpdm(v2g_handle h, const double *x, double *tpdm)
{
    API_BEGIN
    if (!h || !x || !tpdm) throw std::invalid_argument("null argument");
    const v2g::Layout &L = *h->L;
    if (L.nirrep > 1)   // python_api.cc:400-402
        throw std::invalid_argument("not sure how CIWavefunction stores TPDM with symmetry");
    HIPCHK(hipStreamSynchronize(h->st));
    const int64_t a = L.amo, gab = L.gems_ab[0], gaa = L.gems_aa[0];
    const auto ab = block_to_host(h, x, v2g::XB_D2AB, 0, gab * gab);
    const auto aa = block_to_host(h, x, v2g::XB_D2AA, 0, gaa * gaa);
    const auto bb = block_to_host(h, x, v2g::XB_D2BB, 0, gaa * gaa);
    auto GAB = [&](int64_t p, int64_t q) { return (int64_t)L.gab[p * a + q]; };
    auto GAA = [&](int64_t p, int64_t q) { return (int64_t)L.gaa[p * a + q]; };
    // spin-summed <a+_i a+_j a_l a_k> from the ab block (both orderings) and,
    // for i != j, k != l, the aa and bb blocks with their pair signs
    auto d2 = [&](int64_t i, int64_t j, int64_t k, int64_t l) {
        double v = ab[GAB(i, j) * gab + GAB(k, l)] + ab[GAB(j, i) * gab + GAB(l, k)];
        if (i != j && k != l) {
            const double sg = ((i > j) ? -1.0 : 1.0) * ((k > l) ? -1.0 : 1.0);
            const int64_t e = GAA(i, j) * gaa + GAA(k, l);
            v += sg * (aa[e] + bb[e]);
        }
        return v;
    };
    for (int64_t i = 0; i < a; i++)
        for (int64_t j = 0; j < a; j++)
            for (int64_t k = 0; k < a; k++)
                for (int64_t l = 0; l < a; l++) {
                    // python_api.cc:412-452: tpdm[ik][jl] = (D(ij,kl) + D(il,kj)) / 2,
                    // the second term with its D2ab pair in (il,kj) / (li,jk) order
                    double ikjl = d2(i, j, k, l);
                    double iklj = ab[GAB(i, l) * gab + GAB(k, j)] + ab[GAB(l, i) * gab + GAB(j, k)];
                    if (i != l && k != j) {
                        const double sg = ((i > l) ? -1.0 : 1.0) * ((k > j) ? -1.0 : 1.0);
                        const int64_t e = GAA(i, l) * gaa + GAA(k, j);
                        ikjl += sg * (aa[e] + bb[e]);
                    }
                    tpdm[(i * a + k) * (a * a) + j * a + l] = 0.5 * (ikjl + iklj);
                }
    API_END
}

int v2g_spatial_density_dims(v2g_handle h, long long *d1dim, long long *d2dim)
{
    API_BEGIN
    if (!h) throw std::invalid_argument("null handle");
    if (d1dim) *d1dim = v2g::spatial_d1_dim(*h->L);
    if (d2dim) *d2dim = v2g::spatial_d2_dim(*h->L);
    API_END
}

int v2g_singlet_geminals(v2g_handle h, int *counts, int *pairs)
{
    API_BEGIN
    if (!h) throw std::invalid_argument("null handle");
    const v2g::Layout &L = *h->L;
    int64_t o = 0;
    for (int hh = 0; hh < L.nirrep; hh++) {
        if (counts) counts[hh] = L.gems_00[hh];
        for (int n = 0; n < L.gems_00[hh]; n++, o++) {
            const int pr = L.b00[L.b00_off[hh] + n];
            if (pairs) {
                pairs[2 * o] = pr & 0xffff;
                pairs[2 * o + 1] = pr >> 16;
            }
        }
    }
    API_END
}

int v2g_pack_spatial_density(v2g_handle h, const double *x, double *d1, double *d2)
{
    API_BEGIN
    if (!h || !x || !d1 || !d2) throw std::invalid_argument("null argument");
    const v2g::Layout &L = *h->L;
    // the D2 and D1 blocks lead x (dimensions_ order): download up to the end of D1b
    int64_t end = 0;
    for (int hh = 0; hh < L.nirrep; hh++)
        end = std::max(end, L.xoff[v2g::XB_D1B][hh] + (int64_t)L.opt.amopi[hh] * L.opt.amopi[hh]);
    std::vector<double> xh((size_t)end);
    HIPCHK(hipStreamSynchronize(h->st));
    if (end) HIPCHK(hipMemcpy(xh.data(), x, sizeof(double) * end, hipMemcpyDeviceToHost));
    v2g::pack_spatial_density(L, xh.data(), d1, d2);
    API_END
}

int v2g_write_fcidump_rdm(v2g_handle h, const double *x, const char *path)
{
    API_BEGIN
    if (!h || !x || !path) throw std::invalid_argument("null argument");
    HIPCHK(hipStreamSynchronize(h->st));
    const v2g::Layout &L = *h->L;
    const int64_t a = L.amo;
    FILE *f = std::fopen(path, "wb");
    if (!f) throw std::invalid_argument(std::string("cannot open ") + path);
    std::unique_ptr<FILE, int (*)(FILE *)> guard(f, std::fclose);
    auto rec = [&](double v, int p, int q, int r, int s) {
        const int idx[4] = {p, q, r, s};
        if (std::fwrite(&v, sizeof(double), 1, f) != 1 || std::fwrite(idx, sizeof(int), 4, f) != 4)
            throw std::runtime_error("short write");
    };
    // two-electron part (fcidump.cc:183-230): per irrep, every ab geminal pair;
    // active orbitals are numbered from 1 in Pitzer (irrep-major) order
    for (int hh = 0; hh < L.nirrep; hh++) {
        const int64_t g = L.gems_ab[hh], ga = L.gems_aa[hh];
        const auto ab = block_to_host(h, x, v2g::XB_D2AB, hh, g * g);
        const auto aa = block_to_host(h, x, v2g::XB_D2AA, hh, ga * ga);
        const auto bb = block_to_host(h, x, v2g::XB_D2BB, hh, ga * ga);
        const int *bp = L.bab.data() + L.bab_off[hh];
        for (int64_t ij = 0; ij < g; ij++) {
            const int i = bp[ij] & 0xffff, j = bp[ij] >> 16;
            const int64_t ji = L.gab[(int64_t)j * a + i];
            for (int64_t kl = 0; kl < g; kl++) {
                const int k = bp[kl] & 0xffff, l = bp[kl] >> 16;
                const int64_t lk = L.gab[(int64_t)l * a + k];
                double v = ab[ij * g + kl] + ab[ji * g + lk];
                if (i != j && k != l) {
                    const int64_t e = (int64_t)L.gaa[(int64_t)i * a + j] * ga + L.gaa[(int64_t)k * a + l];
                    const double sg = ((i > j) ? -1.0 : 1.0) * ((k > l) ? -1.0 : 1.0);
                    v += sg * (aa[e] + bb[e]);
                }
                rec(v, i + 1, j + 1, k + 1, l + 1);
            }
        }
    }
    // one-electron part (fcidump.cc:238-268): totally symmetric ab geminals
    const int64_t g0 = L.gems_ab[0];
    const int *bp0 = L.bab.data() + L.bab_off[0];
    std::vector<std::vector<double>> d1a(L.nirrep), d1b(L.nirrep);
    for (int hh = 0; hh < L.nirrep; hh++) {
        const int64_t n = L.opt.amopi[hh];
        d1a[hh] = block_to_host(h, x, v2g::XB_D1A, hh, n * n);
        d1b[hh] = block_to_host(h, x, v2g::XB_D1B, hh, n * n);
    }
    for (int64_t ij = 0; ij < g0; ij++) {
        const int i = bp0[ij] & 0xffff, j = bp0[ij] >> 16;
        const int hi = L.sym[i];
        if (hi != L.sym[j]) continue;
        const int64_t n = L.opt.amopi[hi], ii = i - L.poff[hi], jj = j - L.poff[hi];
        rec(d1a[hi][ii * n + jj] + d1b[hi][ii * n + jj], i + 1, j + 1, 0, 0);
    }
    API_END
}

int v2g_mlp_num_params(void) { return v2g::kMlpParams; }

int v2g_mlp_init(v2g_handle h, unsigned long long seed)
{
    API_BEGIN
    if (!h) throw std::invalid_argument("null handle");
    h->ensure_mlp();
    h->mlp_init(seed);
    API_END
}

int v2g_mlp_get_weights(v2g_handle h, float *w)
{
    API_BEGIN
    if (!h || !w) throw std::invalid_argument("null argument");
    h->ensure_mlp();
    HIPCHK(hipMemcpyAsync(w, h->W.p, sizeof(float) * v2g::kMlpParams, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    API_END
}

int v2g_mlp_set_weights(v2g_handle h, const float *w)
{
    API_BEGIN
    if (!h || !w) throw std::invalid_argument("null argument");
    h->ensure_mlp();
    HIPCHK(hipMemcpyAsync(h->W.p, w, sizeof(float) * v2g::kMlpParams, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    API_END
}

int v2g_mlp_loss_grad(v2g_handle h, int nb, const float *feat, const float *tgt, const float *wt,
                      double *grad, double *loss)
{
    API_BEGIN
    if (!h || !feat || !wt || !grad || nb <= 0) throw std::invalid_argument("bad argument");
    h->ensure_mlp();
    DevBuf<float> part, lpart;
    part.alloc((size_t)v2g::mlp_grad_blocks(nb) * v2g::kMlpParams);
    lpart.alloc(v2g::mlp_grad_blocks(nb));
    v2g::mlp_grad(nb, feat, tgt, wt, h->W.p, part.p, lpart.p, h->grad.p, h->dloss.p, h->st);
    HIPCHK(hipMemcpyAsync(grad, h->grad.p, sizeof(double) * v2g::kMlpParams,
                          hipMemcpyDeviceToDevice, h->st));
    double l = 0.0;
    HIPCHK(hipMemcpyAsync(&l, h->dloss.p, sizeof(double), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    if (loss) *loss = l;
    API_END
}

int v2g_mlp_spectral_grad(v2g_handle h, const double *rhs, const double *y, const double *rhi,
                          const double *rlo, float *dout, double *objective)
{
    API_BEGIN
    if (!h || !rhs || !y || !rhi || !rlo || !dout) throw std::invalid_argument("null argument");
    h->ensure_mlp();
    const int64_t n = h->L->ncon;
    if (n > (int64_t)1 << 30) throw std::invalid_argument("too many rows for one batch");
    hipStream_t st = h->st;
    HIPCHK(hipMemcpyAsync(h->yprev.p, y, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    h->mlp_infer(rhs, h->yprev.p);
    auto rho_terms = [&](const double *r, double *kdv, double &a, double &bq) {
        v2g::vec_mul(n, h->lminv.p, r, h->vtmp.p, st);
        v2g::launch_atu(h->plan, h->vtmp.p, h->t.p, st);
        v2g::launch_au(h->plan, h->t.p, kdv, st);
        a = h->dot(r, h->vtmp.p, n);
        bq = h->dot(h->vtmp.p, kdv, n);
    };
    double ab[4];
    rho_terms(rhi, h->kd.p, ab[0], ab[1]);
    rho_terms(rlo, h->kd2.p, ab[2], ab[3]);
    HIPCHK(hipMemcpyAsync(h->scales.p + 2, ab, sizeof(ab), hipMemcpyHostToDevice, st));
    std::vector<int64_t> idx(n);
    for (int64_t i = 0; i < n; i++) idx[i] = i;
    DevBuf<int64_t> d_idx;
    DevBuf<float> feat;
    d_idx.upload(idx);
    feat.alloc((size_t)n * v2g::kMlpIn);
    v2g::mlp_batch_spec((int)n, d_idx.p, h->code.p, h->rownorm.p, rhs, h->yprev.p, rhi, h->kd.p, rlo,
                        h->kd2.p, h->lminv.p, h->scales.p, feat.p, dout, 3.0, 1.0, h->mlp_jacobi_base,
                        0.0, st);
    HIPCHK(hipStreamSynchronize(st));
    if (objective) *objective = std::log(ab[1] / ab[0]) - std::log(ab[3] / ab[2]);
    API_END
}

int v2g_mlp_adam_step(v2g_handle h, const double *grad, double lr)
{
    API_BEGIN
    if (!h || !grad) throw std::invalid_argument("null argument");
    h->ensure_mlp();
    v2g::mlp_adam(h->W.p, grad, h->am.p, h->av.p, lr, ++h->adam_t, h->st);
    HIPCHK(hipStreamSynchronize(h->st));
    API_END
}

int v2g_mlp_minv(v2g_handle h, const double *rhs, const double *y, double *minv)
{
    API_BEGIN
    if (!h || !rhs || !y || !minv) throw std::invalid_argument("null argument");
    h->ensure_mlp();
    h->mlp_infer(rhs, y);
    HIPCHK(hipMemcpyAsync(minv, h->lminv.p, sizeof(double) * h->L->ncon, hipMemcpyDeviceToDevice,
                          h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    API_END
}

const char *v2g_last_error(void) { return v2g::last_error().c_str(); }
const char *v2g_version(void) { return "v2g 0.1 (gfx950)"; }

}  // extern "C"

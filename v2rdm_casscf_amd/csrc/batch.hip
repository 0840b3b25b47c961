// batch.hip -- many independent BPSDP problems (bond-scan points) solved
// together on one GPU.
//
// A batch concatenates the problems: x, z, c of all problems form one primal
// vector, y, b one constraint vector, each problem at its own offset.  Every
// kernel of the iteration covers the whole batch in one launch:
//   * maps: the segment tables of all problems are concatenated; a segment
//     names its problem, whose layout (shifted to the problem's offsets) the
//     row rules read;
//   * eigen split: the blocks of all problems in one list, largest first
//     (size-sorted), each with its problem's mu;
//   * CG: the grid is (reduction blocks, problems); every problem keeps its
//     own scalars (alpha, beta, r.z, iteration, done flag), so problems that
//     converge early stop while the others go on, inside the same launches;
//     the chunk of CG iterations is captured once as a HIP graph (one host
//     thread, one stream) and replayed;
//   * the outer loop keeps mu, the errors, the tolerance schedule and the
//     stopping rule per problem (v2rdm_solver.cc:1778-1912); a converged
//     problem leaves the batch: its vectors stay fixed and its blocks drop out
//     of the splits.
// Results equal the single-problem engine's: the same rows, the same CG
// recurrence and the same block-Jacobi split per problem.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/v2g.h"
#include "common.hpp"
#include "devlayout.hpp"
#include "eig.hpp"
#include "inputs.hpp"
#include "kernels.hpp"
#include "layout.hpp"

namespace v2g {
namespace {

constexpr int kT = 256;
constexpr int kBRed = 128;          // reduction blocks per problem
constexpr int kBChunk = 8;          // CG iterations between host checks (even)
enum : int { P_PAP = 0, P_RZ = 1, P_RR = 2, P_ROZ = 3, P_N = 4 };

struct ProbRange {
    int64_t r0, nr;    // constraint rows of the problem
    int64_t x0, nx;    // primal entries
    int64_t t0, nt;    // row tiles of the A.x launch (p.Ap partials)
};

__device__ __forceinline__ double block_sum(double v, double *sh)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kT / 64; i++) s += sh[i];
    return s;   // valid in thread 0
}

// every block of a problem sums that problem's kBRed partials in one order
__device__ __forceinline__ double all_sum(const double *__restrict__ part, double *sh)
{
    double v = 0.0;
    for (int i = threadIdx.x; i < kBRed; i += kT) v += part[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < kT / 64; i++) s += sh[i];
    __syncthreads();
    return s;
}

__device__ __forceinline__ double *part_of(double *partial, int p, int arr)
{
    return partial + ((int64_t)p * P_N + arr) * kBRed;
}

#define ROWS_OF(pr) for (int64_t i = pr.r0 + blockIdx.x * (int64_t)kT + threadIdx.x; \
                         i < pr.r0 + pr.nr; i += (int64_t)gridDim.x * kT)

// ---- CG over the batch (cg_solver.cc:125-177 per problem) --------------------
__global__ void __launch_bounds__(kT)
k_bcg_init(const ProbRange *__restrict__ prs, const double *__restrict__ b,
           const double *__restrict__ Ap, double *__restrict__ r, double *__restrict__ p,
           double *__restrict__ partial)
{
    __shared__ double sh[kT / 64];
    const int pb = blockIdx.y;
    const ProbRange pr = prs[pb];
    double rr = 0.0;
    ROWS_OF(pr) {
        const double ri = b[i] - Ap[i];
        r[i] = ri;
        p[i] = ri;
        rr += ri * ri;
    }
    const double s = block_sum(rr, sh);
    if (threadIdx.x == 0) {
        part_of(partial, pb, P_RZ)[blockIdx.x] = s;
        part_of(partial, pb, P_RR)[blockIdx.x] = s;
    }
}

__global__ void __launch_bounds__(kT)
k_bcg_init_fin(double *__restrict__ partial, double *__restrict__ scal, const double *__restrict__ tol,
               const int *__restrict__ active, int maxiter)
{
    __shared__ double sh[kT / 64];
    const int pb = blockIdx.x;
    const double rz = all_sum(part_of(partial, pb, P_RZ), sh);
    const double rr = all_sum(part_of(partial, pb, P_RR), sh);
    if (threadIdx.x == 0) {
        double *s = scal + (int64_t)pb * CG_NSLOTS;
        s[CG_NRM0] = sqrt(rr);
        s[CG_RZ0] = rz;
        s[CG_DONE0] = (maxiter <= 0 || !active[pb]) ? 1.0 : 0.0;
        s[CG_DONE1] = 0.0;
        s[CG_ITER] = 0.0;
        s[CG_IT0] = 0.0;
        s[CG_IT1] = 0.0;
        s[CG_TOL] = tol[pb];
        s[CG_MAXIT] = (double)maxiter;
        s[CG_NRM] = 0.0;
    }
}

// p.Ap partials of each problem from the per-tile partials of the A.x launch
__global__ void __launch_bounds__(kT)
k_bcg_rows(const ProbRange *__restrict__ prs, const double *__restrict__ rowpart,
           double *__restrict__ partial, const double *__restrict__ scal, int parity)
{
    __shared__ double sh[kT / 64];
    const int pb = blockIdx.y;
    if (scal[(int64_t)pb * CG_NSLOTS + CG_DONE0 + parity] != 0.0) return;
    const ProbRange pr = prs[pb];
    const int64_t chunk = (pr.nt + gridDim.x - 1) / gridDim.x;
    const int64_t lo = pr.t0 + blockIdx.x * chunk;
    const int64_t hi = (lo + chunk < pr.t0 + pr.nt) ? lo + chunk : pr.t0 + pr.nt;
    double s = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kT) s += rowpart[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) part_of(partial, pb, P_PAP)[blockIdx.x] = t;
}

__global__ void __launch_bounds__(kT)
k_bcg_step(const ProbRange *__restrict__ prs, double *__restrict__ r, const double *__restrict__ Ap,
           double *__restrict__ partial, double *__restrict__ scal, int parity)
{
    __shared__ double sh[kT / 64];
    const int pb = blockIdx.y;
    double *s = scal + (int64_t)pb * CG_NSLOTS;
    if (s[CG_DONE0 + parity] != 0.0) return;
    const ProbRange pr = prs[pb];
    const double pap = all_sum(part_of(partial, pb, P_PAP), sh);
    const double alpha = s[CG_RZ0 + parity] / pap;
    double rr = 0.0;
    ROWS_OF(pr) {
        const double rn = r[i] - alpha * Ap[i];
        r[i] = rn;
        rr += rn * rn;
    }
    const double t = block_sum(rr, sh);
    if (threadIdx.x == 0) {
        part_of(partial, pb, P_RZ)[blockIdx.x] = t;
        part_of(partial, pb, P_RR)[blockIdx.x] = t;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        s[CG_ALPHA] = alpha;
        s[CG_PAP] = pap;
    }
}

__global__ void __launch_bounds__(kT)
k_bcg_finish(const ProbRange *__restrict__ prs, double *__restrict__ y, double *__restrict__ p,
             const double *__restrict__ r, double *__restrict__ partial, double *__restrict__ scal,
             int parity)
{
    __shared__ double sh[kT / 64];
    const int pb = blockIdx.y;
    double *s = scal + (int64_t)pb * CG_NSLOTS;
    if (s[CG_DONE0 + parity] != 0.0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) s[CG_DONE0 + (parity ^ 1)] = 1.0;
        return;
    }
    const ProbRange pr = prs[pb];
    const double alpha = s[CG_ALPHA];
    const double rz = all_sum(part_of(partial, pb, P_RZ), sh);
    const double rr = all_sum(part_of(partial, pb, P_RR), sh);
    const double nrm = sqrt(rr);
    const double rzold = s[CG_RZ0 + parity];
    const bool conv = nrm < s[CG_TOL];
    const int iter = (int)s[CG_IT0 + parity];
    const int iter_next = iter + 1;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        s[CG_NRM] = nrm;
        if (conv) {
            s[CG_DONE0 + (parity ^ 1)] = 1.0;
            s[CG_ITER] = (double)iter;
        } else {
            s[CG_ITER] = (double)iter_next;
            s[CG_DONE0 + (parity ^ 1)] = (iter_next >= (int)s[CG_MAXIT]) ? 1.0 : 0.0;
            s[CG_RZ0 + (parity ^ 1)] = rz;
        }
        s[CG_IT0 + (parity ^ 1)] = (double)iter_next;
    }
    if (conv) {
        ROWS_OF(pr) y[i] += alpha * p[i];
        return;
    }
    const double beta = rz / rzold;
    ROWS_OF(pr) {
        const double pi = p[i];
        y[i] += alpha * pi;
        p[i] = r[i] + beta * pi;
    }
}

// ---- segmented vector work of the outer loop ---------------------------------
// space 0: constraint rows, 1: primal entries; problems with active[p] == 0
// are left untouched
__device__ __forceinline__ void seg_range(const ProbRange &pr, int space, int64_t &lo, int64_t &n)
{
    lo = space ? pr.x0 : pr.r0;
    n = space ? pr.nx : pr.nr;
}

// out = a_p u + b_p v (v may be null)
__global__ void __launch_bounds__(kT)
k_seg_axpby(const ProbRange *__restrict__ prs, int space, const int *__restrict__ active,
            const double *__restrict__ a, const double *u, const double *__restrict__ b,
            const double *v, double *out)
{
    const int pb = blockIdx.y;
    if (active && !active[pb]) return;
    int64_t lo, n;
    seg_range(prs[pb], space, lo, n);
    const double ap = a[pb], bp = v ? b[pb] : 0.0;
    for (int64_t i = lo + blockIdx.x * (int64_t)kT + threadIdx.x; i < lo + n;
         i += (int64_t)gridDim.x * kT)
        out[i] = ap * u[i] + (v ? bp * v[i] : 0.0);
}

// per-problem partial dot products u . v (kBRed per problem)
__global__ void __launch_bounds__(kT)
k_seg_dot(const ProbRange *__restrict__ prs, int space, const double *__restrict__ u,
          const double *__restrict__ v, double *__restrict__ part)
{
    __shared__ double sh[kT / 64];
    const int pb = blockIdx.y;
    int64_t lo, n;
    seg_range(prs[pb], space, lo, n);
    double s = 0.0;
    for (int64_t i = lo + blockIdx.x * (int64_t)kT + threadIdx.x; i < lo + n;
         i += (int64_t)gridDim.x * kT)
        s += u[i] * v[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) part[(int64_t)pb * kBRed + blockIdx.x] = t;
}

__global__ void __launch_bounds__(kT)
k_seg_reduce(const double *__restrict__ part, double *__restrict__ out)
{
    __shared__ double sh[kT / 64];
    const double s = all_sum(part + (int64_t)blockIdx.x * kBRed, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// y <- 2 y - yh, yh <- old y (first call of a problem: yh <- y)
__global__ void __launch_bounds__(kT)
k_seg_extrapolate(const ProbRange *__restrict__ prs, const int *__restrict__ active,
                  const int *__restrict__ first, double *__restrict__ y, double *__restrict__ yh)
{
    const int pb = blockIdx.y;
    if (!active[pb]) return;
    const ProbRange pr = prs[pb];
    const bool f = first[pb] != 0;
    ROWS_OF(pr) {
        const double yc = y[i];
        if (!f) y[i] = 2.0 * yc - yh[i];
        yh[i] = yc;
    }
}

__global__ void __launch_bounds__(kT)
k_seg_copy(const ProbRange *__restrict__ prs, int space, const int *__restrict__ active,
           const double *__restrict__ src, double *__restrict__ dst)
{
    const int pb = blockIdx.y;
    if (!active[pb]) return;
    int64_t lo, n;
    seg_range(prs[pb], space, lo, n);
    for (int64_t i = lo + blockIdx.x * (int64_t)kT + threadIdx.x; i < lo + n;
         i += (int64_t)gridDim.x * kT)
        dst[i] = src[i];
}

}  // namespace
}  // namespace v2g

using v2g::DevBuf;
using v2g::HipError;

struct v2g_batch {
    int np = 0;
    std::vector<std::unique_ptr<v2g::Layout>> L;
    std::vector<std::unique_ptr<DevBuf<int>>> tables;
    DevBuf<v2g::DevLayout> d_lay;
    DevBuf<v2g::Segment> d_rowsegs, d_xsegs, d_t2segs;
    DevBuf<int64_t> d_rowtiles, d_xtiles, d_t2tiles;
    DevBuf<int2> d_g2idx;
    std::vector<v2g::Segment> rowsegs;   // host copy (aux, prob set)
    v2g::MapPlan plan;
    std::vector<v2g::ProbRange> pr;
    DevBuf<v2g::ProbRange> d_pr;
    int64_t dimx = 0, ncon = 0;
    hipStream_t st = nullptr;
    // CG
    DevBuf<double> r, p, Ap, t, partial, scal, rowpart, tolv;
    // loop
    DevBuf<double> Ax, Bv, ATy, yhist, xprev, red, redout, coefa, coefb, mus;
    DevBuf<int> pact, pfirst, pone;
    // eigen split
    std::vector<v2g::EigBlock> small, large;
    DevBuf<v2g::EigBlock> d_small, d_large;
    DevBuf<double> Mw, Aw, Tw, Vw, Qbuf, tol, frop;
    DevBuf<int> active, counter;
    DevBuf<v2g::GemmDesc> d_g1, d_g2, d_g3;
    int max_nb = 0, max_N = 0, max_pairs = 0;
    int64_t max_tiles = 0;
    bool warm = false, pair_wide = false;
    const double *rec_x = nullptr;
    // CG iterations as a captured graph (fixed arguments: y and the buffers)
    hipGraphExec_t cg_exec = nullptr;
    const double *cg_exec_y = nullptr;

    explicit v2g_batch(const v2g_layout_desc *descs, int n);
    ~v2g_batch()
    {
        if (cg_exec) (void)hipGraphExecDestroy(cg_exec);
        if (st) (void)hipStreamDestroy(st);
    }
    // kBRed blocks per problem, always: the reductions sum exactly that many
    // partials (blocks past a small problem's end contribute zeros)
    dim3 grid_of(int) const { return dim3((unsigned)v2g::kBRed, (unsigned)np); }
    void seg_axpby(int space, const int *act, const std::vector<double> &a, const double *u,
                   const std::vector<double> &b, const double *v, double *out);
    std::vector<double> seg_dot(int space, const double *u, const double *v);
    void cg(const double *rhs, double *y, const std::vector<double> &tol, int maxiter,
            std::vector<int> &iters, std::vector<double> &res);
    void update_xz(const std::vector<double> &mu, const std::vector<int> &act, const double *y,
                   const double *c, double *x, double *z, bool capped, int cross_order);
};

v2g_batch::v2g_batch(const v2g_layout_desc *descs, int n) : np(n)
{
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    std::vector<v2g::DevLayout> dls;
    auto tiles = [](const std::vector<v2g::Segment> &segs, std::vector<int64_t> &ts) {
        ts.assign(1, 0);
        for (const auto &s : segs) ts.push_back(ts.back() + v2g::tiles_for(s));
        return ts.back();
    };
    std::vector<v2g::Segment> xsegs, t2segs;
    int64_t g2rows = 0;
    for (int q = 0; q < n; q++) {
        L.emplace_back(new v2g::Layout(v2g::options_from_desc(descs[q])));
        const v2g::Layout &Lq = *L.back();
        auto up = [&](const std::vector<int> &v) {
            tables.emplace_back(new DevBuf<int>());
            tables.back()->upload(v.empty() ? std::vector<int>{0} : v);
            return (const int *)tables.back()->p;
        };
        // block and row offsets of this problem, shifted to its place in the batch
        const v2g::DevLayout dl = v2g::make_dev_layout(
            Lq, {up(Lq.sym), up(Lq.gab), up(Lq.gaa), up(Lq.bab), up(Lq.baa), up(Lq.baab),
                 up(Lq.baba), up(Lq.baaa), up(Lq.b00)},
            dimx, ncon);
        dls.push_back(dl);
        v2g::ProbRange r{};
        r.r0 = ncon;
        r.nr = Lq.ncon;
        r.x0 = dimx;
        r.nx = Lq.dimx;
        r.t0 = 0;
        for (const auto &s : rowsegs) r.t0 += v2g::tiles_for(s);
        for (auto s : Lq.rowsegs) {
            s.start += ncon;
            s.prob = q;
            const bool g2 = s.kind == v2g::RF_G2AB || s.kind == v2g::RF_G2AA;
            s.aux = g2 && !s.wave ? g2rows : -1;
            if (s.aux >= 0) g2rows += s.count;
            rowsegs.push_back(s);
        }
        r.nt = 0;
        for (const auto &s : Lq.rowsegs) r.nt += v2g::tiles_for(s);
        for (auto s : Lq.xsegs) {
            s.start += dimx;
            if (s.kind == v2g::XS_SELF) s.aux += ncon;
            s.prob = q;
            xsegs.push_back(s);
        }
        for (auto s : Lq.xsegs_t2) {
            s.start += dimx;
            s.aux += ncon;
            s.prob = q;
            t2segs.push_back(s);
        }
        pr.push_back(r);
        // eigen blocks of this problem
        for (size_t b = 0; b < Lq.block_dims.size(); b++) {
            const int nb = Lq.block_dims[b];
            if (nb == 0) continue;
            v2g::EigBlock e{};
            e.n = nb;
            e.xoff = Lq.block_offs[b] + dimx;
            e.prob = q;
            if (nb <= 64) {
                e.N = 64;
                small.push_back(e);
            } else {
                e.N = ((nb + 63) / 64) * 64;
                large.push_back(e);
            }
        }
        dimx += Lq.dimx;
        ncon += Lq.ncon;
    }
    if (dimx >= ((int64_t)1 << 30))
        throw std::invalid_argument("batch too large: more than 2^30 primal entries");
    d_lay.upload(dls);
    std::vector<int64_t> ts;
    plan.dlay = d_lay.p;
    plan.row_tiles = tiles(rowsegs, ts);
    d_rowsegs.upload(rowsegs);
    d_rowtiles.upload(ts);
    plan.rowsegs = d_rowsegs.p;
    plan.row_tile_start = d_rowtiles.p;
    plan.nrowsegs = (int)rowsegs.size();
    plan.x_tiles = tiles(xsegs, ts);
    d_xsegs.upload(xsegs);
    d_xtiles.upload(ts);
    plan.xsegs = d_xsegs.p;
    plan.x_tile_start = d_xtiles.p;
    plan.nxsegs = (int)xsegs.size();
    plan.t2_tiles = tiles(t2segs, ts);
    d_t2segs.upload(t2segs);
    d_t2tiles.upload(ts);
    plan.t2segs = d_t2segs.p;
    plan.t2_tile_start = d_t2tiles.p;
    plan.nt2segs = (int)t2segs.size();
    plan.done_stride = v2g::CG_NSLOTS;
    d_pr.upload(pr);
    // the G2 column table: D1 columns of every problem lie between its D1a
    // start and D1b end; the packed form needs them contiguous per problem
    bool packable = g2rows > 0;
    for (int q = 0; q < n && packable; q++) {
        const v2g::Layout &Lq = *L[q];
        int64_t lo = INT64_MAX, hi = -1;
        for (int h = 0; h < Lq.nirrep; h++) {
            const int64_t nn = (int64_t)Lq.opt.amopi[h] * Lq.opt.amopi[h];
            if (!nn) continue;
            lo = std::min(lo, Lq.xoff[v2g::XB_D1A][h]);
            hi = std::max(hi, Lq.xoff[v2g::XB_D1B][h] + nn);
        }
        for (size_t bi = 0; bi < Lq.block_dims.size(); bi++) {
            const int64_t blo = Lq.block_offs[bi], bhi = blo + (int64_t)Lq.block_dims[bi] * Lq.block_dims[bi];
            bool d1 = false;
            for (int h = 0; h < Lq.nirrep; h++)
                d1 = d1 || blo == Lq.xoff[v2g::XB_D1A][h] || blo == Lq.xoff[v2g::XB_D1B][h];
            if (!d1 && bhi > blo && blo < hi && bhi > lo) packable = false;
        }
    }
    if (packable) {
        d_g2idx.alloc(g2rows);
        plan.g2idx = d_g2idx.p;
        // one build per problem: its D1 column range is its own
        for (int q = 0; q < n; q++) {
            const v2g::Layout &Lq = *L[q];
            int64_t lo = INT64_MAX, hi = -1;
            for (int h = 0; h < Lq.nirrep; h++) {
                const int64_t nn = (int64_t)Lq.opt.amopi[h] * Lq.opt.amopi[h];
                if (!nn) continue;
                lo = std::min(lo, Lq.xoff[v2g::XB_D1A][h] + pr[q].x0);
                hi = std::max(hi, Lq.xoff[v2g::XB_D1B][h] + nn + pr[q].x0);
            }
            std::vector<v2g::Segment> mine(rowsegs.size());
            for (size_t i = 0; i < rowsegs.size(); i++) {
                mine[i] = rowsegs[i];
                if (mine[i].prob != q) mine[i].aux = -1;
            }
            v2g::build_g2_index(plan, mine, lo, hi, d_g2idx.p, st);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(st));
    }
    // buffers
    r.alloc(ncon);
    p.alloc(ncon);
    Ap.alloc(ncon);
    t.alloc(dimx);
    partial.alloc((size_t)np * v2g::P_N * v2g::kBRed);
    scal.alloc((size_t)np * v2g::CG_NSLOTS);
    rowpart.alloc(std::max<int64_t>(1, plan.row_tiles));
    tolv.alloc(np);
    Ax.alloc(ncon);
    Bv.alloc(ncon);
    ATy.alloc(dimx);
    yhist.alloc(ncon);
    xprev.alloc(dimx);
    red.alloc((size_t)np * v2g::kBRed);
    redout.alloc(np);
    coefa.alloc(np);
    coefb.alloc(np);
    mus.alloc(np);
    pact.alloc(np);
    pfirst.alloc(np);
    pone.alloc(np);
    {
        std::vector<int> ones(np, 1);
        HIPCHK(hipMemcpy(pone.p, ones.data(), sizeof(int) * np, hipMemcpyHostToDevice));
    }
    // eigen split: large blocks sorted by size (largest first), one lane;
    // 1024-thread pair workgroups only when some padded block reaches 512
    std::stable_sort(large.begin(), large.end(),
                     [](const v2g::EigBlock &a, const v2g::EigBlock &b) { return a.N > b.N; });
    int64_t woff = 0, qoff = 0, poff = 0;
    for (auto &e : large) {
        const int npair = e.N / 64;
        e.woff = woff;
        e.qoff = qoff;
        e.poff = poff;
        woff += (int64_t)e.N * e.N;
        qoff += (int64_t)npair * 64 * 64;
        poff += npair;
        max_nb = std::max(max_nb, e.N / 32);
        max_N = std::max(max_N, e.N);
        max_pairs = std::max(max_pairs, npair);
        max_tiles = std::max(max_tiles, (int64_t)npair * (npair + 1) / 2 + (int64_t)npair * npair);
    }
    pair_wide = max_N >= 512;
    d_small.upload(small);
    d_large.upload(large);
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
        d_g1.alloc(large.size());
        d_g2.alloc(large.size());
        d_g3.alloc(large.size());
    }
}

void v2g_batch::seg_axpby(int space, const int *act, const std::vector<double> &a, const double *u,
                          const std::vector<double> &b, const double *v, double *out)
{
    HIPCHK(hipMemcpyAsync(coefa.p, a.data(), sizeof(double) * np, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(coefb.p, b.data(), sizeof(double) * np, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(v2g::k_seg_axpby, grid_of(space), dim3(v2g::kT), 0, st, d_pr.p, space, act,
                       coefa.p, u, coefb.p, v, out);
}

std::vector<double> v2g_batch::seg_dot(int space, const double *u, const double *v)
{
    hipLaunchKernelGGL(v2g::k_seg_dot, grid_of(space), dim3(v2g::kT), 0, st, d_pr.p, space, u, v, red.p);
    hipLaunchKernelGGL(v2g::k_seg_reduce, dim3(np), dim3(v2g::kT), 0, st, red.p, redout.p);
    std::vector<double> out(np);
    HIPCHK(hipMemcpyAsync(out.data(), redout.p, sizeof(double) * np, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return out;
}

// CG on every problem whose pact flag is set, tolerance per problem
void v2g_batch::cg(const double *rhs, double *y, const std::vector<double> &tolh, int maxiter,
                   std::vector<int> &iters, std::vector<double> &res)
{
    const dim3 gy = grid_of(0);
    HIPCHK(hipMemcpyAsync(tolv.p, tolh.data(), sizeof(double) * np, hipMemcpyHostToDevice, st));
    v2g::launch_atu(plan, y, t.p, st);
    v2g::launch_au(plan, t.p, Ap.p, st);
    hipLaunchKernelGGL(v2g::k_bcg_init, gy, dim3(v2g::kT), 0, st, d_pr.p, rhs, Ap.p, r.p, p.p,
                       partial.p);
    hipLaunchKernelGGL(v2g::k_bcg_init_fin, dim3(np), dim3(v2g::kT), 0, st, partial.p, scal.p, tolv.p,
                       pact.p, maxiter);
    auto enqueue_chunk = [&]() {
        for (int c = 0; c < v2g::kBChunk; c++) {
            const int par = c & 1;
            const double *done = scal.p + v2g::CG_DONE0 + par;
            v2g::launch_atu(plan, p.p, t.p, st, done);
            v2g::launch_au(plan, t.p, Ap.p, st, done, p.p, rowpart.p);
            hipLaunchKernelGGL(v2g::k_bcg_rows, gy, dim3(v2g::kT), 0, st, d_pr.p, rowpart.p, partial.p,
                               scal.p, par);
            hipLaunchKernelGGL(v2g::k_bcg_step, gy, dim3(v2g::kT), 0, st, d_pr.p, r.p, Ap.p, partial.p,
                               scal.p, par);
            hipLaunchKernelGGL(v2g::k_bcg_finish, gy, dim3(v2g::kT), 0, st, d_pr.p, y, p.p, r.p,
                               partial.p, scal.p, par);
        }
    };
    // one host thread drives the batch on its own stream: the chunk is
    // captured once (it holds kernels only) and replayed
    if (!cg_exec || cg_exec_y != y) {
        if (cg_exec) { (void)hipGraphExecDestroy(cg_exec); cg_exec = nullptr; }
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        enqueue_chunk();
        HIPCHK(hipStreamEndCapture(st, &g));
        const hipError_t e = hipGraphInstantiate(&cg_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(e);
        cg_exec_y = y;
    }
    std::vector<double> hs((size_t)np * v2g::CG_NSLOTS);
    int it = 0;
    for (;;) {
        HIPCHK(hipGraphLaunch(cg_exec, st));
        it += v2g::kBChunk;
        HIPCHK(hipMemcpyAsync(hs.data(), scal.p, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        bool all = true;
        for (int q = 0; q < np && all; q++) all = hs[(size_t)q * v2g::CG_NSLOTS + v2g::CG_DONE0 + (it & 1)] != 0.0;
        if (all || it > maxiter + 2 * v2g::kBChunk) break;
    }
    iters.assign(np, 0);
    res.assign(np, 0.0);
    for (int q = 0; q < np; q++) {
        iters[q] = (int)hs[(size_t)q * v2g::CG_NSLOTS + v2g::CG_ITER];
        res[q] = hs[(size_t)q * v2g::CG_NSLOTS + v2g::CG_NRM];
    }
}

// Update_xz for the problems with act[p] != 0 (v2rdm_solver.cc:3216-3289):
// M = sym(A^T y - c + mu_p x), x = P+(M) / mu_p, z = mu_p x - M
void v2g_batch::update_xz(const std::vector<double> &mu, const std::vector<int> &act,
                          const double *y, const double *c, double *x, double *z, bool capped,
                          int cross_order)
{
    HIPCHK(hipMemcpyAsync(mus.p, mu.data(), sizeof(double) * np, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(pact.p, act.data(), sizeof(int) * np, hipMemcpyHostToDevice, st));
    v2g::EigBatch eb;
    eb.mus = mus.p;
    eb.active = pact.p;
    v2g::launch_atu(plan, y, ATy.p, st);
    v2g::eig_small(d_small.p, (int)small.size(), ATy.p, c, x, z, 1.0, st, eb);
    if (large.empty()) return;
    const int nl = (int)large.size();
    const int cold = warm ? 0 : 1;
    capped = capped && !cold;
    v2g::eig_build_m(d_large.p, nl, max_N, ATy.p, c, x, 1.0, Mw.p, Aw.p, Vw.p, cold, st, eb);
    v2g::eig_tol(d_large.p, nl, Mw.p, tol.p, frop.p, capped ? 0 : 1, st, eb);
    // descriptors of the active problems' blocks
    std::vector<v2g::GemmDesc> g1, g2, g3;
    int maxn = 0;
    for (const auto &b : large) {
        if (!act[b.prob]) continue;
        g1.push_back({Mw.p + b.woff, Vw.p + b.woff, Tw.p + b.woff, b.N, b.N, b.N, b.N, b.N, b.N, 0, 0, 0});
        g2.push_back({Vw.p + b.woff, Tw.p + b.woff, Aw.p + b.woff, b.N, b.N, b.N, b.N, b.N, b.N, 1, 0, 1});
        g3.push_back({Tw.p + b.woff, Tw.p + b.woff, x + b.xoff, b.n, b.n, b.N, b.N, b.N, b.n, 0, 1, 1});
        maxn = std::max(maxn, b.n);
    }
    if (g1.empty()) return;
    HIPCHK(hipMemcpyAsync(d_g1.p, g1.data(), sizeof(v2g::GemmDesc) * g1.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_g2.p, g2.data(), sizeof(v2g::GemmDesc) * g2.size(), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_g3.p, g3.data(), sizeof(v2g::GemmDesc) * g3.size(), hipMemcpyHostToDevice, st));
    if (!cold) {
        v2g::gemm_batched(d_g1.p, (int)g1.size(), max_N, max_N, st);
        v2g::gemm_batched(d_g2.p, (int)g2.size(), max_N, max_N, st);
    }
    const int maxsweeps = capped ? 1 : 40;
    const int inner = capped ? 1 : v2g::kEigFullInner;
    for (int sweep = 0; sweep < maxsweeps; sweep++) {
        HIPCHK(hipMemsetAsync(counter.p, 0, sizeof(int), st));
        for (int rnd = 0; rnd < max_nb - 1; rnd++) {
            const int bip = ((capped ? cross_order : 1) && rnd > 0) ? 1 : 0;
            v2g::eig_pairs(d_large.p, nl, max_pairs, rnd, Aw.p, Qbuf.p, active.p, tol.p, counter.p,
                           inner, pair_wide ? 1 : 0, bip, st, eb);
            v2g::eig_apply(d_large.p, nl, max_tiles, rnd, Aw.p, Vw.p, Qbuf.p, active.p, st, eb);
        }
        int cnt = 0;
        HIPCHK(hipMemcpyAsync(&cnt, counter.p, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (cnt == 0) break;
    }
    v2g::eig_scale(d_large.p, nl, Aw.p, Vw.p, Tw.p, 1.0, st, eb);
    v2g::gemm_batched(d_g3.p, (int)g3.size(), maxn, maxn, st);
    v2g::eig_z(d_large.p, nl, x, Mw.p, z, 1.0, st, eb);
    warm = true;
}

extern "C" {

int v2g_batch_create(const v2g_layout_desc *descs, int nprob, v2g_batch_handle *out)
{
    API_BEGIN
    if (!descs || !out || nprob <= 0) throw std::invalid_argument("bad argument");
    *out = new v2g_batch(descs, nprob);
    API_END
}

int v2g_batch_destroy(v2g_batch_handle h)
{
    API_BEGIN
    delete h;
    API_END
}

int v2g_batch_offsets(v2g_batch_handle h, long long *xoff, long long *yoff)
{
    API_BEGIN
    if (!h || !xoff || !yoff) throw std::invalid_argument("null argument");
    for (int q = 0; q < h->np; q++) {
        xoff[q] = h->pr[q].x0;
        yoff[q] = h->pr[q].r0;
    }
    xoff[h->np] = h->dimx;
    yoff[h->np] = h->ncon;
    API_END
}

int v2g_batch_Au(v2g_batch_handle h, const double *x, double *Ax)
{
    API_BEGIN
    if (!h || !x || !Ax) throw std::invalid_argument("null argument");
    v2g::launch_au(h->plan, x, Ax, h->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));
    API_END
}

int v2g_batch_ATu(v2g_batch_handle h, const double *y, double *ATy)
{
    API_BEGIN
    if (!h || !y || !ATy) throw std::invalid_argument("null argument");
    v2g::launch_atu(h->plan, y, ATy, h->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));
    API_END
}

int v2g_batch_inputs(v2g_batch_handle h, const double *const *h1, const double *const *eri,
                     double *c, double *b, double *x, double *z, double *y)
{
    API_BEGIN
    if (!h || !h1 || !eri || !c || !b || !x || !z || !y) throw std::invalid_argument("null argument");
    v2g::build_b_device(h->d_lay.p, h->d_rowsegs.p, (int)h->rowsegs.size(), h->ncon, b, h->st);
    int64_t total = 0;
    for (int q = 0; q < h->np; q++)
        total = std::max<int64_t>(total, 2 * h->pr[q].nx + h->pr[q].nr);
    const std::vector<uint32_t> starts = v2g::guess_chunk_starts(total);
    DevBuf<uint32_t> d_starts;
    d_starts.upload(starts);
    for (int q = 0; q < h->np; q++) {
        const v2g::Layout &Lq = *h->L[q];
        if (!h1[q] || !eri[q]) throw std::invalid_argument("null integrals");
        int mgab = 0, mgaa = 0, mm = 0;
        for (int hh = 0; hh < Lq.nirrep; hh++) {
            mgab = std::max(mgab, Lq.gems_ab[hh]);
            mgaa = std::max(mgaa, Lq.gems_aa[hh]);
            mm = std::max(mm, Lq.opt.amopi[hh]);
        }
        // the layout of problem q alone, with offsets of the batch: c entries
        // land in its slice; the guess is every problem's own srand(0) sequence
        v2g::build_c_device(h->d_lay.p + q, Lq.nirrep, mgab, mgaa, mm, q == 0 ? h->dimx : 0, h1[q],
                            eri[q], c, h->st);
        v2g::guess_device(d_starts.p, 2 * h->pr[q].nx + h->pr[q].nr, h->pr[q].nx, x + h->pr[q].x0,
                          z + h->pr[q].x0, y + h->pr[q].r0, h->st);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));
    API_END
}

int v2g_batch_bpsdp(v2g_batch_handle h, const v2g_params *prm, const double *c, const double *b,
                    double *x, double *z, double *y, v2g_stats *out)
{
    API_BEGIN
    if (!h || !prm || !c || !b || !x || !z || !y || !out) throw std::invalid_argument("null argument");
    if (prm->precond != 0) throw std::invalid_argument("the batched solver runs the reference's plain CG");
    const int np = h->np;
    hipStream_t st = h->st;
    const double tau = prm->tau;
    const bool capped = prm->eig_warm_sweeps > 0 || prm->eig_inner_sweeps > 0;
    std::vector<double> mu(np, 1.0), ep(np, 0.0), ed(np, 0.0), ecur(np, 0.0), edual(np, 0.0);
    std::vector<int> oiter(np, 0), act(np, 1), conv(np, 0), first(np, 1), nexact(np, 0);
    std::vector<long long> cgtot(np, 0);
    std::vector<v2g_stats> s(np);
    const double t0 = v2g::now();
    double tcg = 0.0, txz = 0.0;
    const std::vector<double> ones(np, 1.0), mones(np, -1.0);
    auto upload_act = [&](const std::vector<int> &a) {
        HIPCHK(hipMemcpyAsync(h->pact.p, a.data(), sizeof(int) * np, hipMemcpyHostToDevice, st));
    };
    auto errors = [&](const std::vector<int> &which) {
        // ed = ||A^T y - c + z||, ep = ||A x - b|| per problem (v2rdm_solver.cc:1829-1837)
        upload_act(which);
        v2g::launch_atu(h->plan, y, h->ATy.p, st);
        h->seg_axpby(1, h->pact.p, ones, h->ATy.p, ones, z, h->ATy.p);
        h->seg_axpby(1, h->pact.p, ones, h->ATy.p, mones, c, h->ATy.p);
        const auto edd = h->seg_dot(1, h->ATy.p, h->ATy.p);
        v2g::launch_au(h->plan, x, h->Ax.p, st);
        h->seg_axpby(0, h->pact.p, ones, h->Ax.p, mones, b, h->Ax.p);
        const auto epp = h->seg_dot(0, h->Ax.p, h->Ax.p);
        const auto ec = h->seg_dot(1, c, x);
        const auto ev = h->seg_dot(0, b, y);
        for (int q = 0; q < np; q++)
            if (which[q]) {
                ed[q] = std::sqrt(edd[q]);
                ep[q] = std::sqrt(epp[q]);
                ecur[q] = ec[q];
                edual[q] = ev[q];
            }
    };
    auto passes = [&](int q) {
        return !(ep[q] > prm->r_convergence || ed[q] > prm->r_convergence ||
                 std::fabs(ecur[q] - edual[q]) > prm->e_convergence);
    };
    int nact = np;
    double t_report = t0;
    int bit = 0;
    while (nact > 0) {
        // verbose 2: one progress line every 20 s (long batched solves)
        if (prm->verbose == 2 && v2g::now() - t_report > 20.0) {
            t_report = v2g::now();
            int mx = 0;
            for (int q = 0; q < np; q++) mx = std::max(mx, oiter[q]);
            std::fprintf(stderr, "# batch: %d of %d problems still iterating, %d iterations, "
                         "largest count %d, %.1f s\n", nact, np, bit, mx, t_report - t0);
            std::fflush(stderr);
        }
        bit++;
        const double t1 = v2g::now();
        upload_act(act);
        // Ax = -tau mu (A x - b); z = c - z; B = A z + Ax  (v2rdm_solver.cc:1784-1794)
        std::vector<double> a1(np), b1(np);
        for (int q = 0; q < np; q++) {
            a1[q] = -tau * mu[q];
            b1[q] = tau * mu[q];
        }
        v2g::launch_au(h->plan, x, h->Ax.p, st);
        h->seg_axpby(0, h->pact.p, a1, h->Ax.p, b1, b, h->Ax.p);
        h->seg_axpby(1, h->pact.p, ones, c, mones, z, z);
        v2g::launch_au(h->plan, z, h->Bv.p, st);
        h->seg_axpby(0, h->pact.p, ones, h->Bv.p, ones, h->Ax.p, h->Bv.p);
        std::vector<double> cgtol(np, 0.0);
        for (int q = 0; q < np; q++) {   // v2rdm_solver.cc:1797-1805
            double tl = oiter[q] == 0 ? 0.01 : (ep[q] > ed[q] ? 0.01 * ed[q] : 0.01 * ep[q]);
            cgtol[q] = tl < prm->cg_convergence ? prm->cg_convergence : tl;
        }
        if (prm->cg_extrapolate) {
            HIPCHK(hipMemcpyAsync(h->pfirst.p, first.data(), sizeof(int) * np, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(v2g::k_seg_extrapolate, h->grid_of(0), dim3(v2g::kT), 0, st, h->d_pr.p,
                               h->pact.p, h->pfirst.p, y, h->yhist.p);
            std::fill(first.begin(), first.end(), 0);
        }
        std::vector<int> its;
        std::vector<double> res;
        h->cg(h->Bv.p, y, cgtol, prm->cg_maxiter, its, res);
        const double t2 = v2g::now();
        tcg += t2 - t1;
        if (capped) {
            upload_act(act);
            hipLaunchKernelGGL(v2g::k_seg_copy, h->grid_of(1), dim3(v2g::kT), 0, st, h->d_pr.p, 1,
                               h->pact.p, x, h->xprev.p);
        }
        h->update_xz(mu, act, y, c, x, z, capped, prm->eig_pair_order);
        HIPCHK(hipStreamSynchronize(st));
        txz += v2g::now() - t2;
        errors(act);
        // problems whose tracked split passes the test: redo that split to
        // convergence from the same x, then test again
        std::vector<int> redo(np, 0);
        int nredo = 0;
        for (int q = 0; q < np; q++)
            if (act[q] && capped && passes(q) && h->warm) { redo[q] = 1; nredo++; }
        if (nredo) {
            const double t3 = v2g::now();
            upload_act(redo);
            hipLaunchKernelGGL(v2g::k_seg_copy, h->grid_of(1), dim3(v2g::kT), 0, st, h->d_pr.p, 1,
                               h->pact.p, h->xprev.p, x);
            h->update_xz(mu, redo, y, c, x, z, false, 0);
            HIPCHK(hipStreamSynchronize(st));
            txz += v2g::now() - t3;
            errors(redo);
            for (int q = 0; q < np; q++) nexact[q] += redo[q];
        }
        for (int q = 0; q < np; q++) {
            if (!act[q]) continue;
            cgtot[q] += its[q];
            if (oiter[q] % prm->mu_update_frequency == 0 && oiter[q] > 0) mu[q] = mu[q] * ep[q] / ed[q];
            oiter[q]++;
            v2g_stats &sq = s[q];
            sq.oiter = oiter[q];
            sq.cg_iters = cgtot[q];
            sq.energy_primal = ecur[q];
            sq.energy_dual = edual[q];
            sq.ep = ep[q];
            sq.ed = ed[q];
            sq.mu = mu[q];
            sq.exact_splits = nexact[q];
            const bool done = passes(q);
            if (done || oiter[q] >= prm->maxiter ||
                (prm->max_seconds > 0 && v2g::now() - t0 > prm->max_seconds)) {
                act[q] = 0;
                conv[q] = done ? 1 : 0;
                sq.converged = conv[q];
                sq.t_total = v2g::now() - t0;
                nact--;
            }
        }
    }
    for (int q = 0; q < np; q++) {
        s[q].t_cg = tcg;
        s[q].t_xz = txz;
        out[q] = s[q];
    }
    HIPCHK(hipGetLastError());
    API_END
}

}  // extern "C"

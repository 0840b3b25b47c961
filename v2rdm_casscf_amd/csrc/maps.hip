// maps.hip -- the BPSDP constraint maps A.x and A^T.y on CDNA4.
//
// A.x (bpsdp_Au, v2rdm_solver.cc:3020-3065): one launch.  Every constraint
// row is a gather over a handful of primal entries, computed by one thread;
// the rows whose sums run over all orbitals (traces, S^2, D2->D1
// contractions, maximal-spin rows) get one 64-lane wavefront each.
//
// A^T.y (bpsdp_ATu, v2rdm_solver.cc:3114-3158): for the D / Q / G / spin
// rows also a gather, one thread per primal entry, deterministic and free of
// atomics:
//   * "self" blocks (D200, Q1, Q2, G2): each entry is touched by exactly one
//     row with coefficient +-1, laid out identically -> a coalesced copy;
//   * D2ab / D2aa / D2bb entries invert each family's index rule in closed
//     form (the particle-hole G2 maps are permutations per irrep);
//   * D1 entries sum O(amo) rows -> one wavefront per entry.
// With T1, T2 or spin-adapted Q2 / G2 on, a second launch walks those rows:
// each row's own block entry is a self term, and the terms the rows put into
// Q2 / D2 / G2 / D1 / Q1 entries are added with hardware FP64 atomics, so
// A^T.y with any of them is not bitwise reproducible from run to run (it is
// to rounding).
//
// Row formulas follow d2.cc:296-547, spin.cc:274-493, q2.cc:801-950 and
// 260-467, g2.cc:1061-1260 and 235-546, t1.cc:433-809 and t2.cc:1251-1861.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <vector>

#include "devlayout.hpp"
#include "kernels.hpp"
#include "rowterms.hpp"

namespace v2g {

namespace {

using namespace rows;

// tuning knobs (compile-time; the defaults are the measured best at H50)
#ifndef V2G_MAP_ITEMS_PER_THREAD
#define V2G_MAP_ITEMS_PER_THREAD 4
#endif
#ifndef V2G_AU_WAVES_PER_EU
#define V2G_AU_WAVES_PER_EU 1
#endif

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;
constexpr int kPerThread = V2G_MAP_ITEMS_PER_THREAD;   // independent elementwise items per thread
constexpr int kElemTile = kBlock * kPerThread;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

struct Gather {
    const double *__restrict__ x;
    double acc;
    __device__ __forceinline__ void operator()(int64_t col, double coef) { acc += coef * x[col]; }
};

struct Scatter {
    double *__restrict__ out;
    double y;
    __device__ __forceinline__ void operator()(int64_t col, double coef)
    {
        unsafeAtomicAdd(out + col, coef * y);
    }
};

// Workgroups are dealt to the 8 XCDs round-robin (workgroup i runs on XCD
// i % 8).  xcd_tile gives each XCD one contiguous eighth of the tiles, so the
// rows (or primal entries) that one L2 serves are neighbours and their
// permuted gathers land in a compact part of the input.
constexpr int kXcds = 8;

__device__ __forceinline__ int64_t xcd_tile(int64_t i, int64_t nb)
{
    const int64_t x = i % kXcds, q = nb / kXcds, rem = nb % kXcds;
    return x * q + (x < rem ? x : rem) + i / kXcds;
}

// Chunked variant for launches whose segments differ in cost: XCD x runs the
// chunks x, x + 8, x + 16, ... of kChunk consecutive tiles, which keeps
// neighbouring tiles on one L2 and the XCDs' loads even.  The grid is padded
// to a multiple of 8 kChunk; positions past the last tile do nothing.
constexpr int kChunk = 16;

__device__ __forceinline__ int64_t xcd_chunk_tile(int64_t i)
{
    const int64_t x = i % kXcds, l = i / kXcds;
    return ((l / kChunk) * kXcds + x) * kChunk + (l % kChunk);
}

inline int64_t chunk_grid(int64_t ntiles)
{
    const int64_t c = (int64_t)kXcds * kChunk;
    return (ntiles + c - 1) / c * c;
}

__device__ __forceinline__ int find_segment(const int64_t *__restrict__ tile_start, int nseg,
                                            int64_t b)
{
    int lo = 0, hi = nseg;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (tile_start[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

// block sum of one value per thread in a fixed order (wave butterflies, then
// the waves in order); the result is valid in thread 0
__device__ __forceinline__ double block_total(double v, double *red)
{
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kWavesPerBlock; w++) s += red[w];
    __syncthreads();
    return s;
}

// Column indices of the G2 rows (65 % of all rows at large active spaces),
// built once per layout: for the row visited at position pos of a G2
// segment, g2idx[aux + pos] = (D2 column << 1 | coefficient sign, or -1 when
// the row has no D2 term; D1 column or -1).  A.x then reads one coalesced
// int2 instead of walking the pair and geminal tables before its D2 gather.
struct G2Rec {
    int64_t d1lo, d1hi;     // D1a / D1b columns of x
    int64_t n = 0, c2 = -1, c1 = -1;
    __device__ __forceinline__ void operator()(int64_t col, double coef)
    {
        if (n++ == 0) return;                       // the G2 entry itself
        if (col >= d1lo && col < d1hi) c1 = col;    // delta-selected D1 term (+1)
        else c2 = (col << 1) | (coef < 0.0 ? 1 : 0);
    }
};

__global__ void __launch_bounds__(kBlock)
k_g2_index(const DevLayout *__restrict__ Lp, const Segment *__restrict__ segs, int si, int64_t d1lo,
           int64_t d1hi, int2 *__restrict__ idx)
{
    const Segment sg = segs[si];
    const DevLayout &L = Lp[sg.prob];
    for (int64_t pos = blockIdx.x * (int64_t)kBlock + threadIdx.x; pos < sg.count;
         pos += (int64_t)gridDim.x * kBlock) {
        G2Rec f{d1lo, d1hi};
        elem_terms(L, sg, row_of(L, sg, pos), f);
        idx[sg.aux + pos] = make_int2((int)f.c2, (int)f.c1);
    }
}

// fz.r (optional): the CG residual update fused in, r -= alpha (A.x) with the
// tile sums of the new residual's dot products instead of A.x itself
__global__ void __launch_bounds__(kBlock, V2G_AU_WAVES_PER_EU)
k_au(const DevLayout *__restrict__ Lp, const Segment *__restrict__ segs,
     const int64_t *__restrict__ tile_start, int nseg, const double *__restrict__ x,
     double *__restrict__ Ax, const double *__restrict__ done, const double *__restrict__ dotv,
     double *__restrict__ rowpart, const int2 *__restrict__ g2idx, int done_stride, CgFuse fz)
{
    __shared__ double red[kWavesPerBlock];
    const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const int si = find_segment(tile_start, nseg, tile);
    const Segment sg = segs[si];
    // done flag of this segment's problem (a batch keeps one per problem)
    if (done && done[(int64_t)sg.prob * done_stride] != 0.0) return;
    const DevLayout &L = Lp[sg.prob];
    const int64_t t = tile - tile_start[si];
    const double alpha = fz.r ? fz.scal[CG_ALPHA] : 0.0;
    double contrib = 0.0, rz = 0.0, rr = 0.0, roz = 0.0;
    // row i of A.x has the value v
    auto emit = [&](int64_t i, double v) {
        if (fz.r) {
            const double ro = fz.r[i];
            const double rn = ro - alpha * v;
            fz.r[i] = rn;
            const double zn = fz.minv ? fz.minv[i] * rn : rn;
            rz += rn * zn;
            rr += rn * rn;
            roz += ro * zn;
        } else {
            Ax[i] = v;
            if (dotv) contrib += v * dotv[i];
        }
    };
    if (sg.wave) {
        const int64_t item = t * kWavesPerBlock + (threadIdx.x >> 6);
        if (item < sg.count) {
            const int lane = threadIdx.x & 63;
            Gather f{x, 0.0};
            wave_terms(L, sg, item, lane, f);
            const double v = wave_sum(f.acc);
            if (lane == 0) emit(sg.start + item, v);
        }
    } else if (g2idx && sg.aux >= 0) {
        // G2 rows: self entry, D2 term and D1 term from the precomputed columns
        const int64_t sb = sg.kind == RF_G2AA ? L.xoff[XB_G2AA][sg.h]
                                              : L.xoff[sg.sub ? XB_G2BA : XB_G2AB][sg.h];
        double v[kPerThread];
        int64_t row[kPerThread];
        int2 ci[kPerThread];
#pragma unroll
        for (int u = 0; u < kPerThread; u++) {
            const int64_t pos = t * kElemTile + u * kBlock + threadIdx.x;
            row[u] = -1;
            if (pos < sg.count) {
                ci[u] = g2idx[sg.aux + pos];
                row[u] = row_of(L, sg, pos);
            }
        }
#pragma unroll
        for (int u = 0; u < kPerThread; u++) {
            v[u] = 0.0;
            if (row[u] >= 0) {
                double s = -x[sb + row[u]];
                if (ci[u].x >= 0) {
                    const double d2 = x[ci[u].x >> 1];
                    s += (ci[u].x & 1) ? -d2 : d2;
                }
                if (ci[u].y >= 0) s += x[ci[u].y];
                v[u] = s;
            }
        }
#pragma unroll
        for (int u = 0; u < kPerThread; u++)
            if (row[u] >= 0) emit(sg.start + row[u], v[u]);
    } else {
        // kPerThread independent rows per thread: their gathers overlap in flight
        double v[kPerThread];
        int64_t row[kPerThread];
#pragma unroll
        for (int u = 0; u < kPerThread; u++) {
            const int64_t pos = t * kElemTile + u * kBlock + threadIdx.x;
            Gather f{x, 0.0};
            row[u] = pos < sg.count ? row_of(L, sg, pos) : -1;
            if (row[u] >= 0) elem_terms(L, sg, row[u], f);
            v[u] = f.acc;
        }
#pragma unroll
        for (int u = 0; u < kPerThread; u++)
            if (row[u] >= 0) emit(sg.start + row[u], v[u]);
    }
    if (fz.r) {
        // this workgroup's shares of r.z, r.r and r_old.z, in a fixed order
        const double a = block_total(rz, red);
        const double b2 = block_total(rr, red);
        const double c = block_total(roz, red);
        if (threadIdx.x == 0) {
            fz.part[tile] = a;
            fz.part[fz.ntiles + tile] = b2;
            fz.part[2 * fz.ntiles + tile] = c;
        }
    } else if (rowpart) {
        // this workgroup's share of dotv . Ax, in a fixed order
        const double s2 = block_total(contrib, red);
        if (threadIdx.x == 0) rowpart[tile] = s2;
    }
}

// diag(A A^T): squared norm of every constraint row (Jacobi preconditioner,
// features of the learned preconditioner).  Trace rows repeat columns across
// lanes and are given in closed form.
__global__ void __launch_bounds__(kBlock)
k_row_norm2(const DevLayout *__restrict__ Lp, const Segment *__restrict__ segs,
            const int64_t *__restrict__ tile_start, int nseg, double *__restrict__ d)
{
    const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const int si = find_segment(tile_start, nseg, tile);
    const Segment sg = segs[si];
    const DevLayout &L = Lp[sg.prob];
    const int64_t t = tile - tile_start[si];
    if (sg.wave) {
        const int64_t item = t * kWavesPerBlock + (threadIdx.x >> 6);
        if (item >= sg.count) return;
        const int lane = threadIdx.x & 63;
        double v;
        if (sg.kind == RF_TRACE) {
            const double a = L.amo, npair = a * (a - 1.0) / 2.0;
            // ab diagonal: a^2 entries of weight 1 (or 2); aa diagonal: each
            // geminal appears for (i,j) and (j,i) -> weight 2
            if (sg.sub == 0) v = a * a;
            else if (sg.sub < 3) v = 4.0 * npair;
            else v = 4.0 * a * a + 2.0 * 4.0 * npair;
        } else {
            SumSq f;
            wave_terms(L, sg, item, lane, f);
            v = wave_sum(f.acc);
        }
        if (lane == 0) d[sg.start + item] = v;
    } else {
        for (int u = 0; u < kPerThread; u++) {
            const int64_t pos = t * kElemTile + u * kBlock + threadIdx.x;
            if (pos >= sg.count) return;
            const int64_t e = row_of(L, sg, pos);
            Collect f;
            elem_terms(L, sg, e, f);
            d[sg.start + e] = f.norm2();
        }
    }
}

// eight waves per SIMD (57 VGPRs) to hide the gathers' latency.  sqpart
// (optional): sqpart[tile] = sum of the squares of the tile's entries of
// A^T.y, so that p.(A A^T p) = ||A^T p||^2 needs no second read of p
__global__ void __launch_bounds__(kBlock, 8)
k_atu(const DevLayout *__restrict__ Lp, const Segment *__restrict__ segs,
      const int64_t *__restrict__ tile_start, int nseg, const double *__restrict__ y,
      double *__restrict__ ATy, const double *__restrict__ done, int64_t ntiles, int done_stride,
      double *__restrict__ sqpart)
{
    __shared__ double red[kWavesPerBlock];
    // chunked (not contiguous) XCD ranges: the primal segments differ widely in
    // cost (closed-form D2 gathers vs copies), and contiguous per-XCD ranges
    // would leave some XCDs with all of the expensive tiles
    const int64_t tile = xcd_chunk_tile(blockIdx.x);
    if (tile >= ntiles) return;
    const int si = find_segment(tile_start, nseg, tile);
    const Segment sg = segs[si];
    if (done && done[(int64_t)sg.prob * done_stride] != 0.0) return;
    const DevLayout &L = Lp[sg.prob];
    const int64_t t = tile - tile_start[si];
    double sq = 0.0;
    if (sg.wave) {
        const int64_t item = t * kWavesPerBlock + (threadIdx.x >> 6);
        if (item < sg.count) {
            const int lane = threadIdx.x & 63;
            const double v = wave_sum(atu_d1_wave(L, sg, item, lane, y));
            if (lane == 0) {
                ATy[sg.start + item] = v;
                sq = v * v;
            }
        }
    } else {
        double v[kPerThread];
        int64_t ent[kPerThread];
#pragma unroll
        for (int u = 0; u < kPerThread; u++) {
            const int64_t pos = t * kElemTile + u * kBlock + threadIdx.x;
            v[u] = 0.0;
            ent[u] = pos < sg.count ? entry_of(L, sg, pos) : -1;
            if (ent[u] >= 0) {
                switch (sg.kind) {
                case XS_D2AB: v[u] = atu_d2ab(L, sg, ent[u], y); break;
                case XS_D2AA: v[u] = atu_d2aa(L, sg, ent[u], y); break;
                default: v[u] = atu_self(L, sg, ent[u], y); break;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kPerThread; u++)
            if (ent[u] >= 0) {
                ATy[sg.start + ent[u]] = v[u];
                sq += v[u] * v[u];
            }
    }
    if (sqpart) {
        const double s = block_total(sq, red);
        if (threadIdx.x == 0) sqpart[tile] = s;
    }
}

// second-pass rows (T1, T2, spin-adapted Q2 / G2): self entry written, the
// other terms accumulated
__global__ void __launch_bounds__(kBlock)
k_atu_t2(const DevLayout *__restrict__ Lp, const Segment *__restrict__ segs,
         const int64_t *__restrict__ tile_start, int nseg, const double *__restrict__ y,
         double *__restrict__ ATy, const double *__restrict__ done, int done_stride)
{
    const int64_t tile = blockIdx.x;
    const int si = find_segment(tile_start, nseg, tile);
    const Segment sg = segs[si];
    if (done && done[(int64_t)sg.prob * done_stride] != 0.0) return;
    const DevLayout &L = Lp[sg.prob];
    for (int u = 0; u < kPerThread; u++) {
        const int64_t e = (tile - tile_start[si]) * kElemTile + u * kBlock + threadIdx.x;
        if (e >= sg.count) return;
        const double yv = y[sg.aux + e];
        ATy[sg.start + e] = -yv;
        Scatter f{ATy, yv};
        scatter_terms(L, sg.kind, sg.sub, sg.h, e, f, false);
    }
}

}  // namespace

void launch_au(const MapPlan &p, const double *x, double *Ax, hipStream_t st, const double *done,
               const double *dotv, double *rowpart, CgFuse fz)
{
    if (p.row_tiles > 0)
        hipLaunchKernelGGL(k_au, dim3((unsigned)p.row_tiles), dim3(kBlock), 0, st, p.dlay,
                           p.rowsegs, p.row_tile_start, p.nrowsegs, x, Ax, done, dotv, rowpart,
                           p.g2idx, p.done_stride, fz);
}

void build_g2_index(const MapPlan &p, const std::vector<Segment> &rowsegs, int64_t d1lo,
                    int64_t d1hi, int2 *idx, hipStream_t st)
{
    for (size_t si = 0; si < rowsegs.size(); si++) {
        const Segment &s = rowsegs[si];
        if (s.aux < 0 || s.count == 0) continue;
        const unsigned g = (unsigned)std::min<int64_t>(4096, (s.count + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_g2_index, dim3(g), dim3(kBlock), 0, st, p.dlay, p.rowsegs, (int)si, d1lo,
                           d1hi, idx);
    }
}

void launch_row_norm2(const MapPlan &p, double *d, hipStream_t st)
{
    if (p.row_tiles > 0)
        hipLaunchKernelGGL(k_row_norm2, dim3((unsigned)p.row_tiles), dim3(kBlock), 0, st, p.dlay,
                           p.rowsegs, p.row_tile_start, p.nrowsegs, d);
}

void launch_atu(const MapPlan &p, const double *y, double *ATy, hipStream_t st, const double *done,
                double *sqpart)
{
    // the squares are summed as the entries are written: only valid when no
    // second pass adds to them afterwards
    if (sqpart && p.t2_tiles > 0)
        throw std::logic_error("||A^T y||^2 partials need a layout without second-pass rows");
    if (p.x_tiles > 0)
        hipLaunchKernelGGL(k_atu, dim3((unsigned)chunk_grid(p.x_tiles)), dim3(kBlock), 0, st, p.dlay,
                           p.xsegs, p.x_tile_start, p.nxsegs, y, ATy, done, p.x_tiles, p.done_stride,
                           sqpart);
    if (p.t2_tiles > 0)
        hipLaunchKernelGGL(k_atu_t2, dim3((unsigned)p.t2_tiles), dim3(kBlock), 0, st, p.dlay,
                           p.t2segs, p.t2_tile_start, p.nt2segs, y, ATy, done, p.done_stride);
}

int64_t tiles_for(const Segment &s)
{
    const int64_t per = s.wave ? kWavesPerBlock : kElemTile;
    return (s.count + per - 1) / per;
}

}  // namespace v2g

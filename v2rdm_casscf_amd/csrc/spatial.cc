// spatial.cc -- spin-free active-space densities in the packed layout the
// orbital optimizer reads (v2RDMSolver::PackSpatialDensity,
// v2rdm_solver.cc:3664-3773): the hand-off of an orbital-optimization step.
//
// D1: per irrep, the packed lower triangle (p, q >= p) of D1a + D1b.
// D2: per irrep of the singlet geminals, the packed lower triangle over
// pairs of singlet geminals (ik, jl) (packed index max (max + 1) / 2 + min,
// psi4's INDEX); entry (ik, jl) = half the spin-summed 2-RDM
// G_ikjl = sum <a+_i a+_j a_l a_k> summed over the distinct permutations of
// (ik|jl) that the integrals share, so that the two-electron energy is
// sum_entries d2 (ik|jl): the loop sums the images under i <-> k and
// j <-> l and doubles them off the diagonal (ik != jl) for the pair swap.
// The loop order is the reference's and so is its last-write-wins on an
// entry reached from several (ij, kl).
#include <cstring>

#include "layout.hpp"
#include "spatial.hpp"

namespace v2g {

static inline int64_t tri(int64_t a, int64_t b)
{
    return a > b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a;
}

int64_t spatial_d1_dim(const Layout &L)
{
    int64_t d = 0;
    for (int h = 0; h < L.nirrep; h++) d += (int64_t)L.opt.amopi[h] * (L.opt.amopi[h] + 1) / 2;
    return d;
}

int64_t spatial_d2_dim(const Layout &L)
{
    int64_t d = 0;
    for (int h = 0; h < L.nirrep; h++) d += (int64_t)L.gems_00[h] * (L.gems_00[h] + 1) / 2;
    return d;
}

void pack_spatial_density(const Layout &L, const double *x, double *d1, double *d2)
{
    const int amo = L.amo;
    const int *sym = L.sym.data();
    auto GAB = [&](int p, int q) -> int64_t { return L.gab[p * amo + q]; };
    auto GAA = [&](int p, int q) -> int64_t { return L.gaa[p * amo + q]; };
    int64_t off00[kMaxIrrep + 1] = {0};
    for (int h = 0; h < L.nirrep; h++)
        off00[h + 1] = off00[h] + (int64_t)L.gems_00[h] * (L.gems_00[h] + 1) / 2;
    std::memset(d2, 0, sizeof(double) * off00[L.nirrep]);
    for (int h = 0; h < L.nirrep; h++) {
        const int64_t g = L.gems_ab[h], ga = L.gems_aa[h];
        const double *ab = x + L.xoff[XB_D2AB][h];
        const double *aa = ga ? x + L.xoff[XB_D2AA][h] : nullptr;
        const double *bb = ga ? x + L.xoff[XB_D2BB][h] : nullptr;
        for (int64_t ij = 0; ij < g; ij++) {
            const int pij = L.bab[L.bab_off[h] + ij];
            const int i = pij & 0xffff, j = pij >> 16;
            const int64_t ji = GAB(j, i);
            for (int64_t kl = 0; kl < g; kl++) {
                const int pkl = L.bab[L.bab_off[h] + kl];
                const int k = pkl & 0xffff, l = pkl >> 16;
                const int64_t lk = GAB(l, k);
                const int hik = sym[i] ^ sym[k], hkj = sym[k] ^ sym[j];
                const int64_t ik = L.g00[i * amo + k], jl = L.g00[j * amo + l];
                const int64_t id = off00[hik] + tri(ik, jl);
                const int64_t g2 = L.gems_ab[hkj], ga2 = L.gems_aa[hkj];
                const double *ab2 = x + L.xoff[XB_D2AB][hkj];
                const int64_t kj = GAB(k, j), il = GAB(i, l), jk = GAB(j, k), li = GAB(l, i);
                const double nlj = l == j ? 0.0 : 1.0, nik = i == k ? 0.0 : 1.0;
                double v = 0.0;
                v += 0.5 * ab[ij * g + kl];
                v += 0.5 * ab2[kj * g2 + il] * nlj;
                v += 0.5 * ab2[il * g2 + kj] * nik;
                v += 0.5 * ab[kl * g + ij] * nlj * nik;
                v += 0.5 * ab[ji * g + lk];
                v += 0.5 * ab2[jk * g2 + li] * nlj;
                v += 0.5 * ab2[li * g2 + jk] * nik;
                v += 0.5 * ab[lk * g + ji] * nlj * nik;
                if (i != j && k != l) {
                    const double s = ((i < j) == (k < l)) ? 0.5 : -0.5;
                    const int64_t ija = GAA(i, j), kla = GAA(k, l);
                    v += s * aa[ija * ga + kla];
                    v += s * aa[kla * ga + ija] * nlj * nik;
                    v += s * bb[ija * ga + kla];
                    v += s * bb[kla * ga + ija] * nlj * nik;
                }
                if (k != j && i != l) {
                    const double s = ((k < j) == (i < l)) ? 0.5 : -0.5;
                    const double *aa2 = x + L.xoff[XB_D2AA][hkj], *bb2 = x + L.xoff[XB_D2BB][hkj];
                    const int64_t kja = GAA(k, j), ila = GAA(i, l);
                    v += s * aa2[kja * ga2 + ila] * nlj;
                    v += s * aa2[ila * ga2 + kja] * nik;
                    v += s * bb2[kja * ga2 + ila] * nlj;
                    v += s * bb2[ila * ga2 + kja] * nik;
                }
                if (ik != jl) v *= 2.0;
                d2[id] = v;
            }
        }
    }
    int64_t o = 0;
    for (int h = 0; h < L.nirrep; h++) {
        const int n = L.opt.amopi[h];
        const double *a = x + L.xoff[XB_D1A][h], *b = x + L.xoff[XB_D1B][h];
        for (int i = 0; i < n; i++)
            for (int j = i; j < n; j++) d1[o + tri(i, j)] = a[i * n + j] + b[i * n + j];
        o += (int64_t)n * (n + 1) / 2;
    }
}

}  // namespace v2g

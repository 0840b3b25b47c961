// emu.cc -- bounds-checked host execution of the device row code
// (v2rdm_casscf_amd/csrc/rowterms.hpp) for the CPU test suite.
//
// Test infrastructure: it runs the exact functions the HIP kernels run, so
// every primal / constraint index they form is checked on the CPU against
// the buffers, and their results against the oracle, before any GPU launch.
// x reads go through a checking functor; y reads land in NaN-poisoned guard
// regions when out of range, which poisons the result.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "../../v2rdm_casscf_amd/csrc/devlayout.hpp"
#include "../../v2rdm_casscf_amd/csrc/layout.hpp"
#include "../../v2rdm_casscf_amd/csrc/rowterms.hpp"
#include "../../v2rdm_casscf_amd/csrc/spatial.hpp"

using namespace v2g;

struct Emu {
    Layout L;
    DevLayout D{};
    long long violations = 0;
    explicit Emu(const Options &o) : L(o)
    {
        D = make_dev_layout(L, {L.sym.data(), L.gab.data(), L.gaa.data(), L.bab.data(),
                                L.baa.data(), L.baab.data(), L.baba.data(), L.baaa.data(),
                                L.b00.data()});
    }
};

struct CheckedGather {
    const double *x;
    int64_t n;
    long long *viol;
    double acc = 0.0;
    void operator()(int64_t col, double coef)
    {
        if (col < 0 || col >= n) { (*viol)++; return; }
        acc += coef * x[col];
    }
};

struct CheckedScatter {
    double *out;
    int64_t n;
    long long *viol;
    double y;
    void operator()(int64_t col, double coef)
    {
        if (col < 0 || col >= n) { (*viol)++; return; }
        out[col] += coef * y;
    }
};

extern "C" {

void *emu_create(int nirrep, const int *amopi, int na, int nb, int mult, int q2, int g2, int t2,
                 int spin, int sz, int t1, int sa_q2, int sa_g2, const int *eorder)
{
    Options o;
    o.nirrep = nirrep;
    for (int h = 0; h < 8; h++) o.amopi[h] = h < nirrep ? amopi[h] : 0;
    o.nalpha = na;
    o.nbeta = nb;
    o.multiplicity = mult;
    o.q2 = q2;
    o.g2 = g2;
    o.t2 = t2;
    o.t1 = t1;
    o.spin = spin;
    o.sz = sz;
    o.sa_q2 = sa_q2;
    o.sa_g2 = sa_g2;
    if (eorder) {
        int amo = 0;
        for (int h = 0; h < nirrep; h++) amo += o.amopi[h];
        o.eorder.assign(eorder, eorder + amo);
    }
    try {
        return new Emu(o);
    } catch (...) {
        return nullptr;
    }
}

void emu_destroy(void *h) { delete (Emu *)h; }

void emu_dims(void *h, long long *dimx, long long *ncon)
{
    Emu *E = (Emu *)h;
    *dimx = E->L.dimx;
    *ncon = E->L.ncon;
}

long long emu_violations(void *h) { return ((Emu *)h)->violations; }

// A.x with the device row code; also checks segments tile the rows exactly
void emu_Au(void *h, const double *x, double *Ax)
{
    Emu *E = (Emu *)h;
    const int64_t ncon = E->L.ncon;
    std::vector<char> seen(ncon, 0);
    for (const Segment &sg : E->L.rowsegs) {
        for (int64_t pos = 0; pos < sg.count; pos++) {
            // the row the kernels compute at this position (a permutation of
            // the segment; duplicates or gaps count as violations)
            const int64_t e = sg.wave ? pos : rows::row_of(E->D, sg, pos);
            const int64_t r = sg.start + e;
            if (e < 0 || e >= sg.count || r < 0 || r >= ncon || seen[r]) { E->violations++; continue; }
            seen[r] = 1;
            double v = 0.0;
            if (sg.wave) {
                for (int lane = 0; lane < 64; lane++) {
                    CheckedGather f{x, E->L.dimx, &E->violations};
                    rows::wave_terms(E->D, sg, e, lane, f);
                    v += f.acc;
                }
            } else {
                CheckedGather f{x, E->L.dimx, &E->violations};
                rows::elem_terms(E->D, sg, e, f);
                v = f.acc;
            }
            Ax[r] = v;
        }
    }
    for (int64_t r = 0; r < ncon; r++)
        if (!seen[r]) E->violations++;
}

// A^T.y with the device gathers; y must point into a buffer with ncon
// NaN entries on each side
void emu_ATu(void *h, const double *y, double *ATy)
{
    Emu *E = (Emu *)h;
    const int64_t dimx = E->L.dimx;
    std::vector<char> seen(dimx, 0);
    for (const Segment &sg : E->L.xsegs) {
        for (int64_t pos = 0; pos < sg.count; pos++) {
            const int64_t e = sg.wave ? pos : rows::entry_of(E->D, sg, pos);
            const int64_t i = sg.start + e;
            if (e < 0 || e >= sg.count || i < 0 || i >= dimx || seen[i]) { E->violations++; continue; }
            seen[i] = 1;
            double v = 0.0;
            if (sg.wave) {
                for (int lane = 0; lane < 64; lane++) v += rows::atu_d1_wave(E->D, sg, e, lane, y);
            } else if (sg.kind == XS_D2AB) {
                v = rows::atu_d2ab(E->D, sg, e, y);
            } else if (sg.kind == XS_D2AA) {
                v = rows::atu_d2aa(E->D, sg, e, y);
            } else {
                v = rows::atu_self(E->D, sg, e, y);
            }
            ATy[i] = v;
        }
    }
    for (const Segment &sg : E->L.xsegs_t2) {
        for (int64_t e = 0; e < sg.count; e++) {
            const int64_t i = sg.start + e;
            if (i < 0 || i >= dimx || seen[i]) { E->violations++; continue; }
            seen[i] = 1;
            const double yv = y[sg.aux + e];
            ATy[i] = -yv;
            CheckedScatter f{ATy, dimx, &E->violations, yv};
            rows::scatter_terms(E->D, sg.kind, sg.sub, sg.h, e, f, false);
        }
    }
    for (int64_t i = 0; i < dimx; i++)
        if (!seen[i]) E->violations++;
}

// the packed spatial densities of the orbital-optimization hand-off
void emu_spatial_dims(void *h, long long *d1, long long *d2)
{
    Emu *E = (Emu *)h;
    *d1 = spatial_d1_dim(E->L);
    *d2 = spatial_d2_dim(E->L);
}

void emu_pack_spatial_density(void *h, const double *x, double *d1, double *d2)
{
    pack_spatial_density(((Emu *)h)->L, x, d1, d2);
}

void emu_row_norm2(void *h, double *d)
{
    Emu *E = (Emu *)h;
    for (const Segment &sg : E->L.rowsegs) {
        for (int64_t e = 0; e < sg.count; e++) {
            double v = 0.0;
            if (sg.wave) {
                if (sg.kind == RF_TRACE) {
                    const double a = E->L.amo, npair = a * (a - 1.0) / 2.0;
                    if (sg.sub == 0) v = a * a;
                    else if (sg.sub < 3) v = 4.0 * npair;
                    else v = 4.0 * a * a + 2.0 * 4.0 * npair;
                } else {
                    for (int lane = 0; lane < 64; lane++) {
                        rows::SumSq f;
                        rows::wave_terms(E->D, sg, e, lane, f);
                        v += f.acc;
                    }
                }
            } else {
                rows::Collect f;
                rows::elem_terms(E->D, sg, e, f);
                v = f.norm2();
            }
            d[sg.start + e] = v;
        }
    }
}

}  // extern "C"

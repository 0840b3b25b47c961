// layout.cc -- see layout.hpp.
#include "layout.hpp"

#include <algorithm>
#include <cstring>

#include "devlayout.hpp"
#include <stdexcept>

#include "../../include/v2g.h"

namespace v2g {

Options options_from_desc(const v2g_layout_desc &d)
{
    if (d.nirrep != 1 && d.nirrep != 2 && d.nirrep != 4 && d.nirrep != 8)
        throw std::invalid_argument("nirrep must be 1, 2, 4 or 8");
    Options o;
    o.nirrep = d.nirrep;
    for (int h = 0; h < 8; h++) o.amopi[h] = h < d.nirrep ? d.amopi[h] : 0;
    o.nalpha = d.nalpha;
    o.nbeta = d.nbeta;
    o.multiplicity = d.multiplicity;
    o.q2 = d.constrain_q2;
    o.g2 = d.constrain_g2;
    o.t2 = d.constrain_t2;
    o.t1 = d.constrain_t1;
    o.spin = d.constrain_spin;
    o.sz = d.constrain_sz;
    o.sa_q2 = d.spin_adapt_q2 && d.constrain_q2;
    o.sa_g2 = d.spin_adapt_g2 && d.constrain_g2;
    if (d.energy_order) {
        int amo = 0;
        for (int h = 0; h < o.nirrep; h++) amo += o.amopi[h];
        o.eorder.assign(d.energy_order, d.energy_order + amo);
    }
    return o;
}

Layout::Layout(const Options &o) : opt(o)
{
    if (o.nirrep != 1 && o.nirrep != 2 && o.nirrep != 4 && o.nirrep != 8)
        throw std::invalid_argument("nirrep must be 1, 2, 4 or 8");
    opt.sa_q2 = o.sa_q2 && o.q2;
    opt.sa_g2 = o.sa_g2 && o.g2;
    // v2rdm_solver.cc:538-561
    if ((opt.sa_q2 || opt.sa_g2) && (o.t1 || o.t2))
        throw std::invalid_argument("If constraining T1/T2, G2/Q2 cannot currently be spin adapted.");
    if ((opt.sa_q2 || opt.sa_g2) && o.multiplicity > 1)
        throw std::invalid_argument("Q2 / G2 not spin adapted for S = M != 0");
    // the maximal-spin rows of CONSTRAIN_SPIN read the G2ba block
    // (spin.cc:250-268), which spin-adapted G2 does not lay out
    if (opt.sa_g2 && o.spin)
        throw std::invalid_argument("SPIN_ADAPT_G2 needs CONSTRAIN_SPIN false: the maximal-spin "
                                    "rows address the G2ba block, which spin adaptation replaces");
    nirrep = o.nirrep;
    na = o.nalpha;
    nb = o.nbeta;
    singlet = (o.nalpha == o.nbeta);
    for (int h = 0; h < XB_COUNT; h++)
        for (int i = 0; i < kMaxIrrep; i++) xoff[h][i] = -1;
    for (int f = 0; f < RF_COUNT; f++)
        for (int s = 0; s < 6; s++)
            for (int h = 0; h < kMaxIrrep; h++) rowoff[f][s][h] = -1;
    build_tables();
    build_offsets();
    build_rows();
    build_xsegs();
}

void Layout::build_tables()
{
    amo = 0;
    for (int h = 0; h < nirrep; h++) {
        poff[h] = amo;
        for (int i = 0; i < opt.amopi[h]; i++) sym.push_back(h);
        amo += opt.amopi[h];
    }
    if (amo > 255) throw std::invalid_argument("at most 255 active orbitals");
    gab.assign(amo * amo, -1);
    gaa.assign(amo * amo, -1);
    // Geminal n of irrep h: outer orbital loop i, inner j, pair (first=j,
    // second=i); aa geminals keep first > second (basis.cc:439-548).
    for (int h = 0; h < nirrep; h++) {
        bab_off[h] = (int64_t)bab.size();
        baa_off[h] = (int64_t)baa.size();
        int nab = 0, naa = 0;
        for (int i = 0; i < amo; i++)
            for (int j = 0; j < amo; j++) {
                if ((sym[i] ^ sym[j]) != h) continue;
                const int first = j, second = i;
                gab[first * amo + second] = nab++;
                bab.push_back(first | (second << 16));
                if (first <= second) continue;
                gaa[first * amo + second] = naa;
                gaa[second * amo + first] = naa;
                naa++;
                baa.push_back(first | (second << 16));
            }
        gems_ab[h] = nab;
        gems_aa[h] = naa;
    }
    // singlet geminals: pairs in orbital-energy order, ieo outer and
    // jeo <= ieo inner, stored (orbital of ieo, orbital of jeo)
    // (basis.cc:626-651); read by spin-adapted Q2 and by the packed spatial
    // densities of the orbital optimizer
    {
        std::vector<int> eo = opt.eorder;
        if (eo.empty())
            for (int i = 0; i < amo; i++) eo.push_back(i);
        std::vector<int> chk = eo, id(amo);
        std::sort(chk.begin(), chk.end());
        for (int i = 0; i < amo; i++) id[i] = i;
        if (chk != id)
            throw std::invalid_argument("energy_order must be a permutation of the active orbitals");
        std::vector<std::vector<int>> per(nirrep);
        g00.assign(amo * amo, -1);
        for (int ie = 0; ie < amo; ie++)
            for (int je = 0; je <= ie; je++) {
                const int i = eo[ie], j = eo[je], h = sym[i] ^ sym[j];
                g00[i * amo + j] = g00[j * amo + i] = (int)per[h].size();
                per[h].push_back(i | (j << 16));
            }
        for (int h = 0; h < nirrep; h++) {
            b00_off[h] = (int64_t)b00.size();
            gems_00[h] = (int)per[h].size();
            b00.insert(b00.end(), per[h].begin(), per[h].end());
        }
    }
    if (b00.empty()) b00.push_back(0);
    if (opt.t2 || opt.t1) {
        // triplets: i outer, j, k inner; aab keeps i < j, aaa i < j < k
        // (basis.cc:655-757)
        for (int h = 0; h < nirrep; h++) {
            baab_off[h] = (int64_t)baab.size();
            baba_off[h] = (int64_t)baba.size();
            baaa_off[h] = (int64_t)baaa.size();
            int caab = 0, caba = 0, caaa = 0;
            for (int i = 0; i < amo; i++)
                for (int j = 0; j < amo; j++)
                    for (int k = 0; k < amo; k++) {
                        if ((sym[i] ^ sym[j] ^ sym[k]) != h) continue;
                        const int packed = i | (j << 8) | (k << 16);
                        baba.push_back(packed);
                        caba++;
                        if (i >= j) continue;
                        baab.push_back(packed);
                        caab++;
                        if (j >= k) continue;
                        baaa.push_back(packed);
                        caaa++;
                    }
            trip_aab[h] = caab;
            trip_aba[h] = caba;
            trip_aaa[h] = caaa;
        }
    }
    if (bab.empty()) bab.push_back(0);
    if (baa.empty()) baa.push_back(0);
    if (baab.empty()) baab.push_back(0);
    if (baba.empty()) baba.push_back(0);
    if (baaa.empty()) baaa.push_back(0);
}

void Layout::build_offsets()
{
    int64_t o = 0;
    auto blocks = [&](int w, auto dimf) {
        for (int h = 0; h < nirrep; h++) {
            int64_t d = dimf(h);
            xoff[w][h] = o;
            o += d * d;
            block_dims.push_back((int)d);
            block_offs.push_back(xoff[w][h]);
        }
    };
    auto g = [&](int h) { return (int64_t)gems_ab[h]; };
    auto ga = [&](int h) { return (int64_t)gems_aa[h]; };
    auto n = [&](int h) { return (int64_t)opt.amopi[h]; };
    blocks(XB_D2AB, g);
    blocks(XB_D2AA, ga);
    blocks(XB_D2BB, ga);
    if (opt.spin) {
        if (singlet) blocks(XB_D200, g);
        else blocks(XB_D200, [&](int h) { return 2 * g(h); });
    }
    blocks(XB_D1A, n);
    blocks(XB_D1B, n);
    blocks(XB_Q1A, n);
    blocks(XB_Q1B, n);
    if (q2_plain()) {
        blocks(XB_Q2AB, g);
        blocks(XB_Q2AA, ga);
        blocks(XB_Q2BB, ga);
    } else if (opt.q2) {  // v2rdm_solver.cc:597-609
        blocks(XB_Q2S, [&](int h) { return (int64_t)gems_00[h]; });
        blocks(XB_Q2T, ga);
        blocks(XB_Q2TP1, ga);
        blocks(XB_Q2TM1, ga);
    }
    if (g2_plain()) {
        blocks(XB_G2AB, g);
        blocks(XB_G2BA, g);
        blocks(XB_G2AA, [&](int h) { return 2 * g(h); });
    } else if (opt.g2) {  // v2rdm_solver.cc:623-635
        blocks(XB_G2S, g);
        blocks(XB_G2T, g);
        blocks(XB_G2TP1, g);
        blocks(XB_G2TM1, g);
    }
    if (opt.t1) {  // v2rdm_solver.cc:809-825
        auto t3 = [&](int h) { return (int64_t)trip_aaa[h]; };
        auto t = [&](int h) { return (int64_t)trip_aab[h]; };
        blocks(XB_T1AAA, t3);
        blocks(XB_T1BBB, t3);
        blocks(XB_T1AAB, t);
        blocks(XB_T1BBA, t);
    }
    if (opt.t2) {
        auto T = [&](int h) { return (int64_t)trip_aab[h] + trip_aba[h]; };
        auto t = [&](int h) { return (int64_t)trip_aab[h]; };
        blocks(XB_T2AAA, T);
        blocks(XB_T2BBB, T);
        blocks(XB_T2AAB, t);
        blocks(XB_T2BBA, t);
    }
    dimx = o;
}

void Layout::build_rows()
{
    int64_t r = 0;
    auto seg = [&](int fam, int sub, int h, int dim, int64_t count, int wave) {
        rowoff[fam][sub][h] = r;
        if (count > 0) {
            Segment s{};
            s.kind = fam;
            s.sub = sub;
            s.h = h;
            s.dim = dim;
            s.start = r;
            s.count = count;
            s.wave = wave;
            rowsegs.push_back(s);
        }
        r += count;
    };
    // D2 family, d2.cc:296-547
    if (opt.sz) {
        for (int s = 0; s < 3; s++) seg(RF_TRACE, s, 0, 0, 1, 1);
    } else {
        seg(RF_TRACE, 3, 0, 0, 1, 1);
    }
    for (int s = 0; s < 2; s++)
        for (int h = 0; h < nirrep; h++)
            seg(RF_D1Q1, s, h, opt.amopi[h], (int64_t)opt.amopi[h] * opt.amopi[h], 0);
    const int nk = opt.sz ? 4 : 2;
    for (int kk = 0; kk < nk; kk++) {
        const int sub = opt.sz ? kk : 4 + kk;
        for (int h = 0; h < nirrep; h++)
            seg(RF_CONTR, sub, h, opt.amopi[h], (int64_t)opt.amopi[h] * opt.amopi[h], 1);
    }
    // Spin family, spin.cc:274-493
    if (opt.spin) {
        seg(RF_S2, 0, 0, 0, 1, 1);
        if (singlet) {
            for (int h = 0; h < nirrep; h++)
                seg(RF_SPIN_D1, 0, h, opt.amopi[h], (int64_t)opt.amopi[h] * opt.amopi[h], 0);
            for (int h = 0; h < nirrep; h++)
                seg(RF_SPIN_D2, 0, h, gems_aa[h], (int64_t)gems_aa[h] * gems_aa[h], 0);
            for (int s = 0; s < 2; s++)
                for (int h = 0; h < nirrep; h++)
                    seg(RF_SPIN_ANTI, s, h, gems_aa[h], (int64_t)gems_aa[h] * gems_aa[h], 0);
            for (int h = 0; h < nirrep; h++)
                seg(RF_SPIN_D200, 0, h, gems_ab[h], (int64_t)gems_ab[h] * gems_ab[h], 0);
        } else {
            for (int h = 0; h < nirrep; h++)
                seg(RF_SPIN_D200_4, 0, h, gems_ab[h], 4LL * gems_ab[h] * gems_ab[h], 0);
        }
        if (g2_plain()) {
            seg(RF_MAXSPIN, 0, 0, gems_ab[0], gems_ab[0], 1);
            seg(RF_MAXSPIN, 1, 0, gems_ab[0], gems_ab[0], 1);
        }
    }
    // spin-adapted Q2, q2.cc:260-467: Q2s (singlet geminals), Q2t, Q2t(+1),
    // Q2t(-1) (aa geminals)
    if (opt.q2 && opt.sa_q2) {
        for (int h = 0; h < nirrep; h++)
            seg(RF_Q2SA, 0, h, gems_00[h], (int64_t)gems_00[h] * gems_00[h], 0);
        for (int s = 1; s < 4; s++)
            for (int h = 0; h < nirrep; h++)
                seg(RF_Q2SA, s, h, gems_aa[h], (int64_t)gems_aa[h] * gems_aa[h], 0);
    }
    // Q2, q2.cc:801-950
    if (q2_plain()) {
        for (int h = 0; h < nirrep; h++)
            seg(RF_Q2AB, 0, h, gems_ab[h], (int64_t)gems_ab[h] * gems_ab[h], 0);
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < nirrep; h++)
                seg(RF_Q2AA, s, h, gems_aa[h], (int64_t)gems_aa[h] * gems_aa[h], 0);
    }
    // spin-adapted G2, g2.cc:235-546: G2s, G2t, G2t(+1), G2t(-1)
    if (opt.g2 && opt.sa_g2) {
        for (int s = 0; s < 4; s++)
            for (int h = 0; h < nirrep; h++)
                seg(RF_G2SA, s, h, gems_ab[h], (int64_t)gems_ab[h] * gems_ab[h], 0);
    }
    // G2, g2.cc:1061-1260
    if (g2_plain()) {
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < nirrep; h++)
                seg(RF_G2AB, s, h, gems_ab[h], (int64_t)gems_ab[h] * gems_ab[h], 0);
        for (int h = 0; h < nirrep; h++)
            seg(RF_G2AA, 0, h, gems_ab[h], 4LL * gems_ab[h] * gems_ab[h], 0);
    }
    // T1: aab, bba, aaa, bbb; t1.cc:433-809
    if (opt.t1) {
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < nirrep; h++)
                seg(RF_T1AAB, s, h, trip_aab[h], (int64_t)trip_aab[h] * trip_aab[h], 0);
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < nirrep; h++)
                seg(RF_T1AAA, s, h, trip_aaa[h], (int64_t)trip_aaa[h] * trip_aaa[h], 0);
    }
    // T2 slow maps: aab, bba, aaa(+abb), bbb(+baa); t2.cc:1251-1861
    if (opt.t2) {
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < nirrep; h++)
                seg(RF_T2AAB, s, h, trip_aab[h], (int64_t)trip_aab[h] * trip_aab[h], 0);
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < nirrep; h++) {
                int64_t T = (int64_t)trip_aab[h] + trip_aba[h];
                seg(RF_T2BIG, s, h, (int)T, T * T, 0);
            }
    }
    ncon = r;
}

void Layout::build_xsegs()
{
    auto add = [&](std::vector<Segment> &v, int kind, int sub, int h, int dim, int64_t start,
                   int64_t count, int64_t aux, double coef, int wave) {
        if (count <= 0) return;
        Segment s{};
        s.kind = kind;
        s.sub = sub;
        s.h = h;
        s.dim = dim;
        s.start = start;
        s.count = count;
        s.aux = aux;
        s.coef = coef;
        s.wave = wave;
        v.push_back(s);
    };
    for (int h = 0; h < nirrep; h++)
        add(xsegs, XS_D2AB, 0, h, gems_ab[h], xoff[XB_D2AB][h],
            (int64_t)gems_ab[h] * gems_ab[h], 0, 0.0, 0);
    for (int s = 0; s < 2; s++)
        for (int h = 0; h < nirrep; h++)
            add(xsegs, XS_D2AA, s, h, gems_aa[h], xoff[s ? XB_D2BB : XB_D2AA][h],
                (int64_t)gems_aa[h] * gems_aa[h], 0, 0.0, 0);
    if (opt.spin) {
        for (int h = 0; h < nirrep; h++) {
            int64_t g = gems_ab[h];
            if (singlet)
                add(xsegs, XS_SELF, 0, h, (int)g, xoff[XB_D200][h], g * g,
                    rowoff[RF_SPIN_D200][0][h], 1.0, 0);
            else
                add(xsegs, XS_SELF, 0, h, (int)(2 * g), xoff[XB_D200][h], 4 * g * g,
                    rowoff[RF_SPIN_D200_4][0][h], 1.0, 0);
        }
    }
    for (int s = 0; s < 2; s++)
        for (int h = 0; h < nirrep; h++) {
            int64_t n = opt.amopi[h];
            add(xsegs, XS_D1, s, h, (int)n, xoff[s ? XB_D1B : XB_D1A][h], n * n, 0, 0.0, 1);
        }
    for (int s = 0; s < 2; s++)
        for (int h = 0; h < nirrep; h++) {
            int64_t n = opt.amopi[h];
            add(xsegs, XS_SELF, 0, h, (int)n, xoff[s ? XB_Q1B : XB_Q1A][h], n * n,
                rowoff[RF_D1Q1][s][h], 1.0, 0);
        }
    if (q2_plain()) {
        for (int h = 0; h < nirrep; h++) {
            int64_t g = gems_ab[h];
            add(xsegs, XS_SELF, 0, h, (int)g, xoff[XB_Q2AB][h], g * g, rowoff[RF_Q2AB][0][h],
                -1.0, 0);
        }
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < nirrep; h++) {
                int64_t g = gems_aa[h];
                add(xsegs, XS_SELF, 0, h, (int)g, xoff[s ? XB_Q2BB : XB_Q2AA][h], g * g,
                    rowoff[RF_Q2AA][s][h], -1.0, 0);
            }
    }
    if (g2_plain()) {
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < nirrep; h++) {
                int64_t g = gems_ab[h];
                // sub 1: G2ba of irrep 0 also collects the maximal-spin rows
                int sub = (s == 1 && h == 0 && opt.spin) ? 1 : 0;
                add(xsegs, XS_SELF, sub, h, (int)g, xoff[s ? XB_G2BA : XB_G2AB][h], g * g,
                    rowoff[RF_G2AB][s][h], -1.0, 0);
            }
        for (int h = 0; h < nirrep; h++) {
            int64_t g = gems_ab[h];
            add(xsegs, XS_SELF, 0, h, (int)(2 * g), xoff[XB_G2AA][h], 4 * g * g,
                rowoff[RF_G2AA][0][h], -1.0, 0);
        }
    }
    // spin-adapted Q2 / G2 rows in the second pass: the self term into their
    // own block, the D2 / D1 / Q1 terms scattered atomically (those entries'
    // gathers in the first pass leave the spin-adapted rows out)
    if (opt.q2 && opt.sa_q2) {
        static const XBlock qb[4] = {XB_Q2S, XB_Q2T, XB_Q2TP1, XB_Q2TM1};
        for (int s = 0; s < 4; s++)
            for (int h = 0; h < nirrep; h++) {
                int64_t g = s ? gems_aa[h] : gems_00[h];
                add(xsegs_t2, RF_Q2SA, s, h, (int)g, xoff[qb[s]][h], g * g, rowoff[RF_Q2SA][s][h],
                    -1.0, 0);
            }
    }
    if (opt.g2 && opt.sa_g2) {
        static const XBlock gb[4] = {XB_G2S, XB_G2T, XB_G2TP1, XB_G2TM1};
        for (int s = 0; s < 4; s++)
            for (int h = 0; h < nirrep; h++) {
                int64_t g = gems_ab[h];
                add(xsegs_t2, RF_G2SA, s, h, (int)g, xoff[gb[s]][h], g * g, rowoff[RF_G2SA][s][h],
                    -1.0, 0);
            }
    }
    if (opt.t1) {
        // the triplet pass also walks the T1 rows: self term into the T1 block
        // plus the delta-selected Q2 / D2 / G2 / D1 terms scattered atomically
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < nirrep; h++) {
                int64_t t = trip_aab[h];
                add(xsegs_t2, RF_T1AAB, s, h, (int)t, xoff[s ? XB_T1BBA : XB_T1AAB][h], t * t,
                    rowoff[RF_T1AAB][s][h], -1.0, 0);
            }
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < nirrep; h++) {
                int64_t t = trip_aaa[h];
                add(xsegs_t2, RF_T1AAA, s, h, (int)t, xoff[s ? XB_T1BBB : XB_T1AAA][h], t * t,
                    rowoff[RF_T1AAA][s][h], -1.0, 0);
            }
    }
    if (opt.t2) {
        // the T2 pass walks the T2 rows: self term into the T2 block plus the
        // delta-selected D2 / D1 terms scattered atomically
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < nirrep; h++) {
                int64_t t = trip_aab[h];
                add(xsegs_t2, RF_T2AAB, s, h, (int)t, xoff[s ? XB_T2BBA : XB_T2AAB][h], t * t,
                    rowoff[RF_T2AAB][s][h], -1.0, 0);
            }
        for (int s = 0; s < 2; s++)
            for (int h = 0; h < nirrep; h++) {
                int64_t T = (int64_t)trip_aab[h] + trip_aba[h];
                add(xsegs_t2, RF_T2BIG, s, h, (int)T, xoff[s ? XB_T2BBB : XB_T2AAA][h], T * T,
                    rowoff[RF_T2BIG][s][h], -1.0, 0);
            }
    }
}

std::vector<std::pair<std::string, std::pair<int64_t, int64_t>>> Layout::families() const
{
    static const char *names[RF_COUNT] = {
        "trace", "d1q1", "contract", "s2", "spin_d1", "spin_d2", "spin_anti", "spin_d200",
        "spin_d200_4", "maxspin", "q2ab", "q2aa", "g2ab", "g2aa", "t2aab", "t2big",
        "t1aab", "t1aaa", "q2_spin_adapted", "g2_spin_adapted"};
    std::vector<std::pair<std::string, std::pair<int64_t, int64_t>>> out;
    for (const auto &s : rowsegs) {
        std::string nm = names[s.kind];
        if (!out.empty() && out.back().first == nm &&
            out.back().second.first + out.back().second.second == s.start)
            out.back().second.second += s.count;
        else
            out.push_back({nm, {s.start, s.count}});
    }
    return out;
}

DevLayout make_dev_layout(const Layout &L, const DevTables &t, int64_t dx, int64_t dr)
{
    DevLayout dl;
    std::memset(&dl, 0, sizeof(dl));
    dl.nirrep = L.nirrep;
    dl.amo = L.amo;
    dl.singlet = L.singlet;
    dl.sz = L.opt.sz;
    dl.spin = L.opt.spin;
    dl.q2 = L.q2_plain();
    dl.g2 = L.g2_plain();
    dl.t2 = L.opt.t2;
    dl.t1 = L.opt.t1;
    dl.sa_q2 = L.opt.q2 && L.opt.sa_q2;
    dl.sa_g2 = L.opt.g2 && L.opt.sa_g2;
    dl.na = L.na;
    dl.nb = L.nb;
    dl.ms = (L.opt.multiplicity - 1.0) / 2.0;
    for (int h = 0; h < kMaxIrrep; h++) {
        dl.amopi[h] = L.opt.amopi[h];
        dl.poff[h] = L.poff[h];
        dl.gems_ab[h] = L.gems_ab[h];
        dl.gems_aa[h] = L.gems_aa[h];
        dl.gems_00[h] = L.gems_00[h];
        dl.trip_aab[h] = L.trip_aab[h];
        dl.trip_aba[h] = L.trip_aba[h];
        dl.trip_aaa[h] = L.trip_aaa[h];
        dl.bab_off[h] = L.bab_off[h];
        dl.baa_off[h] = L.baa_off[h];
        dl.baab_off[h] = L.baab_off[h];
        dl.baba_off[h] = L.baba_off[h];
        dl.baaa_off[h] = L.baaa_off[h];
        dl.b00_off[h] = L.b00_off[h];
    }
    for (int b = 0; b < XB_COUNT; b++)
        for (int h = 0; h < kMaxIrrep; h++)
            dl.xoff[b][h] = L.xoff[b][h] >= 0 ? L.xoff[b][h] + dx : -1;
    for (int f = 0; f < RF_COUNT; f++)
        for (int s = 0; s < 6; s++)
            for (int h = 0; h < kMaxIrrep; h++)
                dl.rowoff[f][s][h] = L.rowoff[f][s][h] >= 0 ? L.rowoff[f][s][h] + dr : -1;
    dl.sym = t.sym;
    dl.gab = t.gab;
    dl.gaa = t.gaa;
    dl.bab = t.bab;
    dl.baa = t.baa;
    dl.baab = t.baab;
    dl.baba = t.baba;
    dl.baaa = t.baaa;
    dl.b00 = t.b00;
    return dl;
}

}  // namespace v2g

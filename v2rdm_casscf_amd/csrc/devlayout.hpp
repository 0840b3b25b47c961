// devlayout.hpp -- plain-old-data copy of the layout that device code reads
// (one instance in device memory per solver handle; accessed through a
// pointer so every field read is a uniform scalar load).
#pragma once
#include <cstdint>

#include "layout.hpp"

namespace v2g {

struct DevLayout {
    int nirrep, amo, singlet, sz, spin;
    int q2, g2;        // plain (not spin-adapted) Q2 / G2 rows present
    int t2, t1;
    int sa_q2, sa_g2;  // spin-adapted Q2 / G2 rows present
    double na, nb;
    double ms;         // M_s of the state, (multiplicity - 1) / 2
    int amopi[kMaxIrrep], poff[kMaxIrrep];
    int gems_ab[kMaxIrrep], gems_aa[kMaxIrrep];
    int trip_aab[kMaxIrrep], trip_aba[kMaxIrrep], trip_aaa[kMaxIrrep];
    int64_t bab_off[kMaxIrrep], baa_off[kMaxIrrep], baab_off[kMaxIrrep], baba_off[kMaxIrrep];
    int64_t baaa_off[kMaxIrrep];
    int gems_00[kMaxIrrep];
    int64_t b00_off[kMaxIrrep];
    int64_t xoff[XB_COUNT][kMaxIrrep];
    int64_t rowoff[RF_COUNT][6][kMaxIrrep];
    // device tables
    const int *sym;    // amo
    const int *gab;    // amo*amo
    const int *gaa;    // amo*amo
    const int *bab;    // packed pairs, per-irrep offsets bab_off
    const int *baa;
    const int *baab;   // packed triplets
    const int *baba;
    const int *baaa;   // i < j < k triplets (T1aaa / T1bbb)
    const int *b00;    // singlet geminals (spin-adapted Q2), per-irrep offsets b00_off
};

// where a layout's tables live for the code that reads a DevLayout
struct DevTables {
    const int *sym, *gab, *gaa, *bab, *baa, *baab, *baba, *baaa, *b00;
};

// the device view of layout L with its tables at t; block and row offsets
// shifted by dx / dr (a problem's place in a batch, 0 alone)
DevLayout make_dev_layout(const Layout &L, const DevTables &t, int64_t dx = 0, int64_t dr = 0);

}  // namespace v2g

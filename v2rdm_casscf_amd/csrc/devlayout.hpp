// devlayout.hpp -- plain-old-data copy of the layout that device code reads
// (one instance in device memory per solver handle; accessed through a
// pointer so every field read is a uniform scalar load).
#pragma once
#include <cstdint>

#include "layout.hpp"

namespace v2g {

struct DevLayout {
    int nirrep, amo, singlet, sz, spin, q2, g2, t2, t1;
    double na, nb;
    double ms;         // M_s of the state, (multiplicity - 1) / 2
    int amopi[kMaxIrrep], poff[kMaxIrrep];
    int gems_ab[kMaxIrrep], gems_aa[kMaxIrrep];
    int trip_aab[kMaxIrrep], trip_aba[kMaxIrrep], trip_aaa[kMaxIrrep];
    int64_t bab_off[kMaxIrrep], baa_off[kMaxIrrep], baab_off[kMaxIrrep], baba_off[kMaxIrrep];
    int64_t baaa_off[kMaxIrrep];
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
};

}  // namespace v2g

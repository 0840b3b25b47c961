// layout.hpp -- geminal / triplet tables, primal-block offsets and the
// constraint-row segment table of the BPSDP maps, built on the host and
// mirrored to device memory.
//
// Follows v2RDMSolver::BuildBasis (basis.cc:439-548, 655-757) for the
// geminal and triplet numbering and v2rdm_solver.cc:560-1270 for dimx_,
// the x block order (dimensions_) and nconstraints_.  The explicit segment
// table replaces the reference's implicit `offset` cursor
// (v2rdm_solver.h:295): each segment names a contiguous range of rows
// (A.x direction) or of primal entries (A^T.y direction) and the family
// rule that computes it.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

struct v2g_layout_desc;

namespace v2g {

constexpr int kMaxIrrep = 8;

// primal (x) blocks, dimensions_ order of v2rdm_solver.cc:1129-1270
enum XBlock : int {
    XB_D2AB = 0, XB_D2AA, XB_D2BB, XB_D200, XB_D1A, XB_D1B, XB_Q1A, XB_Q1B,
    XB_Q2AB, XB_Q2AA, XB_Q2BB, XB_G2AB, XB_G2BA, XB_G2AA,
    XB_T2AAA, XB_T2BBB, XB_T2AAB, XB_T2BBA,
    XB_T1AAA, XB_T1BBB, XB_T1AAB, XB_T1BBA,
    // spin-adapted Q2 / G2 (SPIN_ADAPT_Q2 / _G2), v2rdm_solver.cc:597-635
    XB_Q2S, XB_Q2T, XB_Q2TP1, XB_Q2TM1, XB_G2S, XB_G2T, XB_G2TP1, XB_G2TM1,
    XB_COUNT
};

// constraint-row families in bpsdp_Au order
enum RowFam : int {
    RF_TRACE = 0,      // sub: 0 ab, 1 aa, 2 bb, 3 total (CONSTRAIN_SZ false)
    RF_D1Q1,           // sub: spin
    RF_CONTR,          // sub: 0 ab->D1a, 1 ab->D1b, 2 aa->D1a, 3 bb->D1b,
                       //      4 (aa+ab)->D1a, 5 (bb+ab)->D1b
    RF_S2,
    RF_SPIN_D1,
    RF_SPIN_D2,        // D2aa - D2bb
    RF_SPIN_ANTI,      // sub: 0 D2aa rows, 1 D2bb rows
    RF_SPIN_D200,      // singlet
    RF_SPIN_D200_4,    // non-singlet 2g x 2g
    RF_MAXSPIN,        // sub: orientation 0 (kl,ii), 1 (ii,kl)
    RF_Q2AB,
    RF_Q2AA,           // sub: spin
    RF_G2AB,           // sub: 0 G2ab, 1 G2ba
    RF_G2AA,           // 2g x 2g super block
    RF_T2AAB,          // sub: 0 T2aab, 1 T2bba
    RF_T2BIG,          // sub: 0 T2aaa(+abb), 1 T2bbb(+baa)
    RF_T1AAB,          // sub: 0 T1aab, 1 T1bba
    RF_T1AAA,          // sub: 0 T1aaa, 1 T1bbb
    RF_Q2SA,           // spin-adapted Q2, sub: 0 Q2s, 1 Q2t, 2 Q2t(+1), 3 Q2t(-1)
    RF_G2SA,           // spin-adapted G2, sub: 0 G2s, 1 G2t, 2 G2t(+1), 3 G2t(-1)
    RF_COUNT
};

// primal-side segment kinds for A^T.y
enum XSegKind : int {
    XS_D2AB = 0,       // gather (closed-form inverse rules)
    XS_D2AA,           // sub: spin; gather
    XS_D1,             // sub: spin; one wavefront per entry
    XS_SELF,           // x[e] = coef * y[row0 + e] (+ max-spin extras)
    XS_COUNT
};

struct Segment {
    int32_t kind;      // RowFam (rows) or XSegKind (primal)
    int32_t sub;
    int32_t h;
    int32_t dim;       // block dimension (g, ga, n, 2g, T ...)
    int64_t start;     // first row / first primal entry
    int64_t count;     // rows / entries
    int64_t aux;       // XS_SELF: first source row; else unused
    double coef;       // XS_SELF: coefficient
    int32_t wave;      // 1: one wavefront per item
    int32_t prob;      // problem of a batch whose layout the segment follows (0 alone)
};

struct Options {
    int nirrep = 1;
    int amopi[kMaxIrrep] = {0};
    int nalpha = 0, nbeta = 0, multiplicity = 1;
    bool q2 = true, g2 = true, t2 = false, spin = true, sz = true, t1 = false;
    bool sa_q2 = false, sa_g2 = false;  // SPIN_ADAPT_Q2 / SPIN_ADAPT_G2
    // active orbitals from lowest to highest orbital energy (order of the
    // singlet geminals of SPIN_ADAPT_Q2); empty = Pitzer order
    std::vector<int> eorder;
};

// the problem description of the C ABI as layout options, with the
// reference's refusals (v2rdm_solver.cc:538-561) as std::invalid_argument
Options options_from_desc(const v2g_layout_desc &d);

class Layout {
public:
    explicit Layout(const Options &o);

    Options opt;
    int nirrep = 1, amo = 0, na = 0, nb = 0;
    bool singlet = true;
    std::vector<int> sym;               // irrep of active orbital p (Pitzer order)
    int poff[kMaxIrrep] = {0};
    std::vector<int> gab, gaa;          // amo x amo: geminal number within its irrep
    int gems_ab[kMaxIrrep] = {0}, gems_aa[kMaxIrrep] = {0};
    std::vector<int> bab, baa;          // packed pairs p | q << 16, concatenated over irreps
    int64_t bab_off[kMaxIrrep] = {0}, baa_off[kMaxIrrep] = {0};
    int trip_aab[kMaxIrrep] = {0}, trip_aba[kMaxIrrep] = {0}, trip_aaa[kMaxIrrep] = {0};
    std::vector<int> baab, baba, baaa;  // packed triplets i | j << 8 | k << 16
    int64_t baab_off[kMaxIrrep] = {0}, baba_off[kMaxIrrep] = {0}, baaa_off[kMaxIrrep] = {0};
    // singlet geminals (SPIN_ADAPT_Q2, packed spatial densities):
    // energy-ordered pairs i | j << 16, and the index of (i, j) in its irrep
    int gems_00[kMaxIrrep] = {0};
    std::vector<int> b00;
    std::vector<int> g00;               // amo x amo, symmetric
    int64_t b00_off[kMaxIrrep] = {0};
    bool q2_plain() const { return opt.q2 && !opt.sa_q2; }
    bool g2_plain() const { return opt.g2 && !opt.sa_g2; }

    int64_t xoff[XB_COUNT][kMaxIrrep];  // -1 when the block is absent
    int64_t dimx = 0, ncon = 0;
    std::vector<int> block_dims;        // dimensions_ order
    std::vector<int64_t> block_offs;

    // row offsets of each family: rowoff[fam][sub][h]
    int64_t rowoff[RF_COUNT][6][kMaxIrrep];
    std::vector<Segment> rowsegs;       // A.x segments, ascending start
    std::vector<Segment> xsegs;         // A^T.y segments (main pass)
    std::vector<Segment> xsegs_t2;      // A^T.y triplet pass (T1 / T2 self + scatter)

    std::vector<std::pair<std::string, std::pair<int64_t, int64_t>>> families() const;

private:
    void build_tables();
    void build_offsets();
    void build_rows();
    void build_xsegs();
};

}  // namespace v2g

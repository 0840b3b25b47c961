// inputs.hpp -- the BPSDP path inputs (b, c, the RANDOM guess) built on the
// device (inputs.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "devlayout.hpp"

namespace v2g {

// outputs of the glibc generator produced by one thread of the guess kernel
constexpr int64_t kGuessChunk = 65536;

// BuildConstraints: b over all ncon rows (of every problem of a batch plan)
void build_b_device(const DevLayout *dlay, const Segment *rowsegs, int nseg, int64_t ncon, double *b,
                    hipStream_t st);
// RepackIntegrals without core orbitals: c over dimx entries
void build_c_device(const DevLayout *dlay, int nirrep, int max_gab, int max_gaa, int max_m,
                    int64_t dimx, const double *h1, const double *eri, double *c, hipStream_t st);
// glibc state (31 words, oldest first) after srand(seed), before the first rand()
void glibc_rand_state_after_srand(unsigned seed, uint32_t (&state)[31]);
// the generator's state at the start of every chunk of kGuessChunk outputs
// of the sequence after srand(0), 31 words per chunk
std::vector<uint32_t> guess_chunk_starts(int64_t total);
// x, z (interleaved) then y from total = 2 dimx + ncon outputs
void guess_device(const uint32_t *starts, int64_t total, int64_t dimx, double *x, double *z, double *y,
                  hipStream_t st);

}  // namespace v2g

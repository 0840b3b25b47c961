"""The device guess restates glibc's rand() (TYPE_3 additive feedback
generator) and jumps ahead with powers of its 31 x 31 step matrix (mod 2^32).
Both are pinned here on the CPU against the C library's own rand() after
srand(0); the device kernel is held to the host guess bit for bit by the GPU
tests."""
import ctypes
import ctypes.util

import numpy as np

M32 = (1 << 32) - 1


def glibc_sequence(n, seed=0):
    """rand() outputs after srand(seed): r_i = 16807 r_(i-1) mod (2^31 - 1)
    for i < 31 (seed 0 acts as 1), r_31..33 = r_0..2, r_i = r_(i-31) + r_(i-3)
    mod 2^32, outputs r_(i) >> 1 from i = 344 on."""
    r = [seed or 1]
    for i in range(1, 31):
        r.append((16807 * r[-1]) % 2147483647)
    r += r[0:3]
    while len(r) < 344 + n:
        r.append((r[-31] + r[-3]) & M32)
    return [v >> 1 for v in r[344:344 + n]], r


def test_restatement_matches_the_c_library():
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.srand(0)
    ref = [libc.rand() for _ in range(5000)]
    got, _ = glibc_sequence(5000)
    assert got == ref


def test_jump_ahead_matrix_gives_chunk_start_states():
    chunk, nchunks = 256, 6
    _, r = glibc_sequence(chunk * nchunks)
    T = np.zeros((31, 31), dtype=np.uint64)
    for i in range(30):
        T[i, i + 1] = 1
    T[30, 0] = T[30, 28] = 1

    def mul(A, B):
        return (A.astype(object).dot(B.astype(object)) % (1 << 32)).astype(np.uint64)

    P = T.copy()
    e = 1
    while e < chunk:
        P = mul(P, P)
        e *= 2
    s = np.array(r[344 - 31:344], dtype=np.uint64)
    for c in range(nchunks):
        t0 = 344 + c * chunk
        assert list(s) == r[t0 - 31:t0]
        s = (P.astype(object).dot(s.astype(object)) % (1 << 32)).astype(np.uint64)

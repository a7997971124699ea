"""The measurement kernel's pair enumeration, replayed on the CPU.

eval_pairs (tdvmc_amd/csrc/kernels_eval.hip) visits the pairs of a walker in
64 x 64 tiles of particle blocks: lane l holds row particle 64 A + l, the
column particles ride a ring so that at step s lane l meets ring slot
(l + s) mod 64; diagonal tiles take s = 1 .. 32 (s = 32 by lanes < 32 only);
a last block of at most 32 particles is held twice in the ring when full
blocks precede it (folded: slot q holds particle q mod 32, 32 steps); the
diagonal tile of a last block of r <= 32 particles takes s = 1 .. r - 1; and
the steps of all tiles are split evenly over the walker's waves.  This test
repeats that arithmetic (the same counts, offsets and validity tests) and
checks that every unordered pair occurs exactly once, that no particle meets
itself, and that the step count adds up, for many particle counts and 1, 2
or 4 waves per walker.  The GPU parity tests check the kernel's sums.
"""
import pytest


def _enumerate(N, NW):
    nfull, rem = N // 64, N % 64
    nblk = nfull + (1 if rem else 0)
    fold = 0 < rem <= 32 and nfull > 0
    dlast = rem - 1 if 0 < rem <= 32 else 32
    total = nfull * (nfull - 1) // 2 * 64 + nfull * 32 + ((dlast + nfull * (32 if fold else 64)) if rem else 0)
    seen = {}
    for wave in range(NW):
        g0, g1 = total * wave // NW, total * (wave + 1) // NW
        off = 0
        for A in range(nblk):
            for Bk in range(A, nblk):
                diag = A == Bk
                folded = (not diag) and fold and Bk == nblk - 1
                cnt = (dlast if (rem > 0 and A == nblk - 1) else 32) if diag else (32 if folded else 64)
                lo, hi = max(g0, off), min(g1, off + cnt)
                off += cnt
                if lo >= hi:
                    continue
                s_first = lo - (off - cnt) + (1 if diag else 0)
                cmask = 31 if folded else 63
                for u in range(hi - lo):
                    s = s_first + u
                    for lane in range(64):
                        ra = A * 64 + lane
                        ci = Bk * 64 + ((lane + s) & cmask)
                        if ra < N and ci < N and (not diag or s < 32 or lane < 32):
                            assert ra != ci, (N, A, Bk, s, lane)
                            key = (min(ra, ci), max(ra, ci))
                            seen[key] = seen.get(key, 0) + 1
        assert off == total
    return seen


@pytest.mark.parametrize("NW", [1, 2, 4])
def test_every_pair_once(NW):
    for N in list(range(2, 200)) + [343, 400, 511, 512, 513]:
        seen = _enumerate(N, NW)
        assert len(seen) == N * (N - 1) // 2, N
        assert all(v == 1 for v in seen.values()), N

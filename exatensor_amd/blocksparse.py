"""Seeded block-sparse tensor pairs for the batched contraction path.

Stands in for ExaTENSOR's hierarchical tensors (INTRAVIRT/tensor_recursive.F90,
TensContractionSplit :9202-9800), which produce the lists of block
contractions that the block-level caller then runs one per call
(INTERVIRT/tavp_worker.F90:9440-9480).

Pattern D(a,b,c,d) += L(a,k,c,l) * R(l,b,k,d): every mode is split into S
segments with extents drawn from a seeded set; each L and R block is present
with probability `fill`; every (a,b,c,d,k,l) segment combination with both
blocks present is one block contraction into D block (a,b,c,d).
"""
import random

PATTERN = "D(a,b,c,d)+=L(a,k,c,l)*R(l,b,k,d)"
EXTENTS = (8, 12, 16, 24, 32, 48, 64)


def generate(seed=2026, segments=7, fill=0.2, max_ops=4096, extents=EXTENTS):
    rng = random.Random(seed)
    ext = {m: [rng.choice(extents) for _ in range(segments)] for m in "abcdkl"}
    S = range(segments)
    lblocks = {(a, k, c, l) for a in S for k in S for c in S for l in S if rng.random() < fill}
    rblocks = {(l, b, k, d) for l in S for b in S for k in S for d in S if rng.random() < fill}
    ops = []
    for (a, k, c, l) in sorted(lblocks):
        for b in S:
            for d in S:
                if (l, b, k, d) in rblocks:
                    ops.append(((a, b, c, d), (a, k, c, l), (l, b, k, d)))
    if max_ops is not None and len(ops) > max_ops:
        ops = sorted(rng.sample(ops, max_ops))
    return ext, ops


def dims(ext, key, modes):
    return [ext[m][s] for m, s in zip(modes, key)]


def flops(ext, ops):
    tot = 0
    for dk, lk, rk in ops:
        a, b, c, d = dims(ext, dk, "abcd")
        k, l = ext["k"][lk[1]], ext["l"][lk[3]]
        tot += 2 * a * b * c * d * k * l
    return tot

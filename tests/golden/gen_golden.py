"""Writes the golden fixtures under tests/golden/ from the reference TAL-SH CPU
library (CP-TAL + talsh.h C API, -DNO_GPU -DNO_BLAS), compiled from its own
sources by oracle/ref.mk into oracle/_ref/libtalsh_ref.so.

Run in the development container (where the reference sources exist):
    make -f oracle/ref.mk && python tests/golden/gen_golden.py
Fixtures hold inputs and the reference's outputs only.
"""
import ctypes as C
import gzip
import json
import os
import random
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_lib import C4, C8, NP_DTYPE, R4, R8, REF_SO, Oracle, Reference  # noqa: E402

REF_TREE = "/root/reference/TALSH"


def shape_list_patterns():
    out = []
    for fn in ("tensor_contractions.txt", "tensor_contractions48.txt", "tensor_contractions_problem.txt"):
        p = os.path.join(REF_TREE, fn)
        if not os.path.exists(p):
            continue
        lines = [ln.rstrip("\n") for ln in open(p)]
        for i in range(0, len(lines) - 1, 2):
            pat = lines[i].strip()
            nums = [int(x) for x in lines[i + 1].split()]
            dr = nums[0]; dd = nums[1:1 + dr]
            lr = nums[1 + dr]; ld = nums[2 + dr:2 + dr + lr]
            rr = nums[2 + dr + lr]; rd = nums[3 + dr + lr:3 + dr + lr + rr]
            out.append((pat, dd, ld, rd))
    return out


def random_pattern(rng, max_rank=5, conj=True, hyper=False):
    """A random valid contraction pattern string with shuffled index order."""
    ncon = rng.randint(0, 3)
    nl = rng.randint(0, 3)
    nr = rng.randint(0, 3)
    nh = rng.randint(0, 1) if hyper else 0
    labels = [f"x{i}" if rng.random() < 0.3 else chr(ord('a') + i) for i in range(ncon + nl + nr + nh)]
    con = labels[:ncon]; lu = labels[ncon:ncon + nl]; ru = labels[ncon + nl:ncon + nl + nr]
    hy = labels[ncon + nl + nr:]
    dl = lu + ru + hy
    rng.shuffle(dl)
    ll = con + lu + hy
    rl = con + ru + hy
    rng.shuffle(ll); rng.shuffle(rl)
    cj = lambda: "+" if (conj and rng.random() < 0.2) else ""
    sep = lambda: "|" if rng.random() < 0.1 else ","
    def t(name, idx):
        s = name + cj() + "("
        for k, x in enumerate(idx):
            s += (sep() if k else "") + x
        return s + ")"
    return t("D", dl) + "+=" + t("L", ll) + "*" + t("R", rl)


def main():
    ref = Reference()
    orc = Oracle()
    rng = random.Random(20260117)
    # ------------------------------------------------------------------ patterns
    pats = [
        "D(a,b,c,d)+=L(d,i,b,j)*R(j,c,i,a)", "D(a,b,i,j)+=L(a,k,i,l)*R(l,b,k,j)",
        "D(a,b,c,d)+=L(a,k,c,l)*R(l,b,k,d)", "D(a,b,i,j)+=L(c,b,d,a)*R(j,d,i,c)",
        "D(a,b,c,d)+=L(c,i,j,a)*R(b,j,d,i)", "D(a,b,c,d)=L(c,d,b,a)", "D(a,b,c,d)+=L(c,d,b,a)",
        "D(ia1,ib2)+=L(ib2,k2+,c3)*R+(c3+,ia1,k2)", "D(a,b)+=L+(b,a)", "D+(a,b)+=L(a,i)*R(i,b)",
        "D()+=L(a,b)*R(b,a)", "D(a,b)+=L(a,b)*R()", "D(a,b)+=L()*R(b,a)", "D()+=L()*R()",
        "D(a,b)+=L(a,b)*2.0", "D(a,b)=L(b,a)*0.5", "D(a,b,c)+=L(a,c)*R(c,b,a)",
        "D(a|b)+=L(a|i)*R(i|b)", "D(a,a)=L(a)", "D(a)=L(b)", "D(a,b)+=L(a,b,b)*R(x)", "D(a,)+=L(a)",
        "D(a,b)+=L(a,i)*R(i,b)*R(b)", "D(a,b+)+=L(c,a)*R(b|c+)", "D(+a,b)+=L(a,b)", "D(a b)+=L(a,b)",
        "D(a,b)+=L(a,i", "", "*", "D(a)*", "D(a,b)+=L(i,a)*R(i,b,i)", "D(a,b)+=L(i,a)*R(j,b)",
        "D(a)+=L(a,i)*R(i,j)", "D(i,j)+=L(i,k)*R(k,j)", "D(a,b,c,d,e,f)=L(f,e,d,c,b,a)",
        "D(A1,B2)+=L(B2,Q)*R(Q,A1)", "D(_a)+=L(_a)", "D(a)+=L_+(a)", "D(a,b)+=L(a,i)*R+(i,b)",
        "D(a,b)+=L(i,a)*R+(b,i)", "D+(a,b)+=L(i,a)*R(b,i)", "D(a,b)+=L(a,b)*R(a)", "D(a,b)+=L(a)*R(a,b)",
        "D(a,b,c)+=L(c,a)*R(b)", "X(a,b)+=Y(b,a)*Z()", "D(a,b)+=L(b,a)*R()*R()",
    ]
    shape_pats = shape_list_patterns()
    pats += [p for p, *_ in shape_pats]
    pats += [random_pattern(rng, hyper=(i % 4 == 0)) for i in range(300)]
    # malformed mutations of valid patterns
    for i in range(120):
        p = random_pattern(rng)
        k = rng.randrange(len(p))
        op = rng.randrange(3)
        if op == 0:
            p = p[:k] + p[k + 1:]
        elif op == 1:
            p = p[:k] + rng.choice("(),|+*= x1") + p[k:]
        else:
            p = p[:k] + rng.choice("(),|+*=") + p[k + 1:]
        pats.append(p)
    seen = set(); uniq = []
    for p in pats:
        if p not in seen:
            seen.add(p); uniq.append(p)
    pattern_rows = []
    for p in uniq:
        e, dig, dr, lr, rr, cj = ref.str2dig(p)
        pattern_rows.append(dict(pattern=p, ierr=e, dig=dig, drank=dr, lrank=lr, rrank=rr, conj=cj))
    # ------------------------------------------------------------ permutations
    perm_rows = []
    for row in pattern_rows:
        if row["ierr"] != 0 or row["rrank"] < 0:
            continue
        for hyper in (0, 1):
            for tl, tr in ((1, 0), (0, 0), (1, 1), (0, 1)):
                for cj in sorted({row["conj"], 0, 2, 4, 1}):
                    res = ref.permutations(hyper, tl, tr, row["lrank"], row["rrank"], row["dig"], cj)
                    perm_rows.append(dict(hyper=hyper, tl=tl, tr=tr, lrank=row["lrank"], rrank=row["rrank"],
                                          dig=row["dig"], conj=cj, **res))
    with open(os.path.join(HERE, "patterns.json"), "w") as f:
        json.dump(dict(source="reference TAL-SH CPU library (talsh_get_contr_ptrn_str2dig)",
                       rows=pattern_rows), f, separators=(",", ":"))
    with gzip.open(os.path.join(HERE, "permutations.json.gz"), "wt") as f:
        json.dump(dict(source="reference get_contr_permutations / get_contraction_permutations",
                       rows=perm_rows), f, separators=(",", ":"))
    # ------------------------------------------------------------ tensor ops
    arrays = {}
    meta = []
    nrs = np.random.default_rng(7)

    def rnd(kind, n):
        if kind in (C4, C8):
            return (nrs.uniform(-1, 1, n) + 1j * nrs.uniform(-1, 1, n)).astype(NP_DTYPE[kind])
        return nrs.uniform(-1, 1, n).astype(NP_DTYPE[kind])

    # copies: random permutations, all kinds, ranks 1..6, small extents
    for i in range(48):
        kind = (R4, R8, C4, C8)[i % 4]
        rank = 1 + (i % 6)
        dims = [int(nrs.integers(1, 6)) for _ in range(rank)]
        perm = list(nrs.permutation(rank))
        lab = [chr(ord('a') + k) for k in range(rank)]
        conj = "+" if (kind in (C4, C8) and i % 8 >= 4) else ""
        # L(lab) -> D(lab permuted): D position of L index k is perm[k]
        dlab = [None] * rank
        for k in range(rank):
            dlab[perm[k]] = lab[k]
        pat = "D(" + ",".join(dlab) + ")=L" + conj + "(" + ",".join(lab) + ")"
        ddims = [dims[lab.index(x)] for x in dlab]
        l = rnd(kind, int(np.prod(dims)))
        e, out = ref.copy(pat, ddims, dims, l)
        key = f"copy{i}"
        arrays[key + "_l"] = l; arrays[key + "_d"] = out
        meta.append(dict(op="copy", key=key, kind=kind, pattern=pat, ldims=dims, ddims=ddims, ierr=e))
    # adds
    for i in range(16):
        kind = (R4, R8, C4, C8)[i % 4]
        rank = 1 + (i % 4)
        dims = [int(nrs.integers(1, 6)) for _ in range(rank)]
        perm = list(nrs.permutation(rank))
        lab = [chr(ord('a') + k) for k in range(rank)]
        dlab = [None] * rank
        for k in range(rank):
            dlab[perm[k]] = lab[k]
        conj = "+" if (kind in (C4, C8) and i % 8 >= 4) else ""
        pat = "D(" + ",".join(dlab) + ")+=L" + conj + "(" + ",".join(lab) + ")"
        ddims = [dims[lab.index(x)] for x in dlab]
        l = rnd(kind, int(np.prod(dims))); d = rnd(kind, int(np.prod(dims)))
        scale = complex(0.75, -0.5) if kind in (C4, C8) else 0.75
        e, out = ref.add(pat, d, ddims, l, dims, scale)
        key = f"add{i}"
        arrays[key + "_l"] = l; arrays[key + "_d0"] = d; arrays[key + "_d"] = out
        meta.append(dict(op="add", key=key, kind=kind, pattern=pat, ldims=dims, ddims=ddims, ierr=e,
                         scale=[scale.real, scale.imag] if isinstance(scale, complex) else [scale, 0.0]))
    # init / scale
    for i, kind in enumerate((R4, R8, C4, C8)):
        dims = [3, 4, 5]
        d = rnd(kind, 60)
        v = complex(1.25, -2.5)
        e1, o1 = ref.init(d, dims, v)
        e2, o2 = ref.scale(d, dims, v)
        key = f"elt{i}"
        arrays[key + "_d0"] = d; arrays[key + "_init"] = o1; arrays[key + "_scale"] = o2
        meta.append(dict(op="init_scale", key=key, kind=kind, ddims=dims, value=[v.real, v.imag],
                         ierr=[e1, e2]))
    # contractions: the configured patterns at reduced extents, the benchmark
    # shape lists at reduced extents, and random patterns; all data kinds.
    cases = [("D(a,b,c,d)+=L(d,i,b,j)*R(j,c,i,a)", 5), ("D(a,b,i,j)+=L(a,k,i,l)*R(l,b,k,j)", 4),
             ("D(a,b,c,d)+=L(a,k,c,l)*R(l,b,k,d)", 3), ("D(a,b,i,j)+=L(c,b,d,a)*R(j,d,i,c)", 4),
             ("D()+=L(a,b)*R(b,a)", 6), ("D(a,b)+=L(b,a)*R()", 5), ("D(a,b)+=L()*R(b,a)", 5),
             ("D()+=L()*R()", 1), ("D(a,b,c)+=L(c,a)*R(b)", 4), ("D(a,b)+=L(a,i)*R(i,b)", 7)]
    for p, *_ in shape_pats[::3]:
        cases.append((p, 3))
    for i in range(24):
        cases.append((random_pattern(rng, conj=False), 3))
    for ci, (p, ext) in enumerate(cases):
        e, dig, dr, lr, rr, cj = ref.str2dig(p)
        if e or rr < 0:
            continue
        # extents per label: varied, small
        edim = {}
        def ext_of(label_key):
            if label_key not in edim:
                edim[label_key] = int(nrs.integers(max(1, ext - 2), ext + 2))
            return edim[label_key]
        ldims = [0] * lr; rdims = [0] * rr; ddims = [0] * dr
        for k in range(lr):
            v = dig[k]
            x = ext_of(("d", v) if v > 0 else ("c", k))
            ldims[k] = x
            if v > 0:
                ddims[v - 1] = x
            else:
                rdims[-v - 1] = x
        for k in range(rr):
            v = dig[lr + k]
            if v > 0:
                if ddims[v - 1] == 0:
                    ddims[v - 1] = ext_of(("d", v))
                rdims[k] = ddims[v - 1]
        for kind in (R8, R4, C8, C4):
            for acc in (True, False):
                if (ci + kind + acc) % 2 and ci > 10:
                    continue
                vd = int(np.prod(ddims)) if dr else 1
                vl = int(np.prod(ldims)) if lr else 1
                vr = int(np.prod(rdims)) if rr else 1
                d = rnd(kind, vd); l = rnd(kind, vl); r = rnd(kind, vr)
                alpha = complex(0.5, 0.25) if kind in (C4, C8) else 1.5
                e, out = ref.contract(p, d, ddims, l, ldims, r, rdims, alpha, acc)
                key = f"ctr{ci}_{kind}_{int(acc)}"
                arrays[key + "_d0"] = d; arrays[key + "_l"] = l; arrays[key + "_r"] = r
                if out is not None:
                    arrays[key + "_d"] = out
                meta.append(dict(op="contract", key=key, kind=kind, pattern=p, ddims=ddims, ldims=ldims,
                                 rdims=rdims, alpha=[alpha.real, alpha.imag] if isinstance(alpha, complex)
                                 else [alpha, 0.0], accumulative=acc, ierr=e))
    np.savez_compressed(os.path.join(HERE, "tensor_ops.npz"), **arrays)
    with open(os.path.join(HERE, "tensor_ops.json"), "w") as f:
        json.dump(dict(source="reference TAL-SH CPU library, DEV_HOST ops", cases=meta), f, indent=0)
    print(f"patterns {len(pattern_rows)} permutations {len(perm_rows)} ops {len(meta)}")


def slice_insert_fixtures(ref):
    """talshTensorSlice / talshTensorInsert on the host: random boxes of every
    data kind, ranks 1..5, accumulative and not, full-extent boxes included."""
    nrs = np.random.default_rng(11)
    arrays, meta = {}, []

    def rnd(kind, n):
        if kind in (C4, C8):
            return (nrs.uniform(-1, 1, n) + 1j * nrs.uniform(-1, 1, n)).astype(NP_DTYPE[kind])
        return nrs.uniform(-1, 1, n).astype(NP_DTYPE[kind])

    for i in range(40):
        kind = (R4, R8, C4, C8)[i % 4]
        rank = 1 + (i % 5)
        tdims = [int(nrs.integers(1, 7)) for _ in range(rank)]
        full = i % 10 == 9
        sdims = list(tdims) if full else [int(nrs.integers(1, t + 1)) for t in tdims]
        offs = [0] * rank if full else [int(nrs.integers(0, t - s + 1)) for t, s in zip(tdims, sdims)]
        acc = bool(i % 3 == 1)
        tens = rnd(kind, int(np.prod(tdims)))
        slc = rnd(kind, int(np.prod(sdims)))
        e1, o1 = ref.slice(tens, tdims, slc, sdims, offs, acc)
        e2, o2 = ref.insert(tens, tdims, slc, sdims, offs, acc)
        key = f"box{i}"
        arrays[key + "_t"] = tens; arrays[key + "_s"] = slc
        if o1 is not None:
            arrays[key + "_slice"] = o1
        if o2 is not None:
            arrays[key + "_insert"] = o2
        meta.append(dict(key=key, kind=kind, tdims=tdims, sdims=sdims, offsets=offs, accumulative=acc,
                         ierr=[e1, e2]))
    np.savez_compressed(os.path.join(HERE, "slice_insert.npz"), **arrays)
    with open(os.path.join(HERE, "slice_insert.json"), "w") as f:
        json.dump(dict(source="reference TAL-SH CPU library, talshTensorSlice/Insert on DEV_HOST", cases=meta), f,
                  indent=0)
    print(f"slice/insert cases {len(meta)}")


def decompose_fixtures(ref):
    """Leaves of the reference's recursive talshTensorOpDecompose2 halving of
    contractions (batch, right, left and contracted splits)."""
    import tensop_ctypes as TC
    rng = random.Random(77)
    cases = [("D(a,b,c,d)+=L(d,i,b,j)*R(j,c,i,a)", [9, 7, 8, 5], 3), ("D(a,b,i,j)+=L(a,k,i,l)*R(l,b,k,j)", [6, 5, 7, 4], 4),
             ("D(a,b)+=L(a,i)*R(i,b)", [37, 5], 60), ("D(a,b)+=L(i,a)*R(i,b)", [3, 4], 200),
             ("D(a,b,c)+=L(a,c,i)*R(b,i,c)", [10, 11, 6], 13), ("D()+=L(a,b)*R(b,a)", [], 0)]
    rows = []
    for ci, (pat, _, _) in enumerate(cases):
        e, dig, dr, lr, rr, cj = ref.str2dig(pat)
        for trial in range(4):
            ext = {}
            lab_l = pat[pat.index("L(") + 2:pat.index(")", pat.index("L("))].split(",")
            lab_r = pat[pat.index("R(") + 2:pat.index(")", pat.index("R("))].split(",")
            lab_d = pat[2:pat.index(")")].split(",") if dr else []
            for x in lab_l + lab_r + lab_d:
                if x and x not in ext:
                    ext[x] = rng.randint(1, 23)
            dd = [ext[x] for x in lab_d]
            ld = [ext[x] for x in lab_l]
            rd = [ext[x] for x in lab_r]
            vol = sum(int(np.prod(v)) if v else 1 for v in (dd, ld, rd))
            limit = max(3, vol // rng.choice((2, 5, 17, 60)))
            leaves = TC.decompose(C.CDLL(REF_SO), pat, R8, dd, ld, rd, limit)
            rows.append(dict(pattern=pat, ddims=dd, ldims=ld, rdims=rd, limit=limit, leaves=leaves))
    with open(os.path.join(HERE, "decompose.json"), "w") as f:
        json.dump(dict(source="reference TAL-SH talshTensorOpDecompose2 (recursive halving)", cases=rows), f,
                  separators=(",", ":"))
    print(f"decompose cases {len(rows)} leaves {sum(len(r['leaves']) for r in rows)}")


if __name__ == "__main__":
    if "--slice-insert" in sys.argv:
        slice_insert_fixtures(Reference())
    elif "--decompose" in sys.argv:
        decompose_fixtures(Reference())
    else:
        main()
        slice_insert_fixtures(Reference())
        decompose_fixtures(Reference())

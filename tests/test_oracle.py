"""Pins the CP-TAL oracle (oracle/cptal.c).

Checks, in order of strength:
  * golden fixtures written by the reference TAL-SH CPU library (tests/golden);
  * the live reference library, when it has been built (randomised cases);
  * closed forms from the reference's own tests (TALSH/test.cpp) and
    brute-force / long-double evaluation on small inputs.
"""
import gzip
import json
import os
import random

import numpy as np
import pytest

from oracle_lib import C4, C8, NP_DTYPE, R4, R8

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    p = os.path.join(GOLD, name)
    if name.endswith(".gz"):
        with gzip.open(p, "rt") as f:
            return json.load(f)
    with open(p) as f:
        return json.load(f)


# ----------------------------------------------------------------- patterns
def test_patterns_match_reference_fixtures(oracle):
    rows = _load("patterns.json")["rows"]
    assert len(rows) > 400
    nbad = 0
    for row in rows:
        e, dig, dr, lr, rr, cj = oracle.str2dig(row["pattern"])
        assert e == row["ierr"], row
        if e == 0:
            assert (dig, dr, lr, rr, cj) == (row["dig"], row["drank"], row["lrank"], row["rrank"], row["conj"]), row
        else:
            nbad += 1
    assert nbad > 50  # malformed inputs are covered too


def test_pattern_known_answers(oracle):
    # configured workloads (digital forms derived by hand)
    assert oracle.str2dig("D(a,b,c,d)+=L(d,i,b,j)*R(j,c,i,a)")[:2] == (0, [4, -3, 2, -1, -4, 3, -2, 1])
    assert oracle.str2dig("D(a,b,i,j)+=L(a,k,i,l)*R(l,b,k,j)")[:2] == (0, [1, -3, 3, -1, -4, 2, -2, 4])
    # unary copy gets a fake rank-0 R
    e, dig, dr, lr, rr, cj = oracle.str2dig("D(a,b,c,d)=L(c,d,b,a)")
    assert (e, dig, dr, lr, rr) == (0, [3, 4, 2, 1], 4, 4, 0)
    assert oracle.str2dig("D(a,b)+=L+(a,i)*R+(i,b)")[5] == 6


def test_pattern_roundtrip_sym(oracle):
    rows = _load("patterns.json")["rows"]
    for row in rows:
        if row["ierr"] or row["rrank"] < 0:
            continue
        if any(v > 0 for v in row["dig"]) and len(set(v for v in row["dig"] if v > 0)) != sum(
                1 for v in row["dig"] if v > 0):
            continue  # hyper-index patterns do not round-trip through the symbolic printer
        e, s = oracle.dig2str(row["lrank"], row["rrank"], row["conj"], row["dig"])
        assert e == 0
        e2, dig2, dr2, lr2, rr2, cj2 = oracle.str2dig(s)
        assert e2 == 0 and dig2 == row["dig"] and cj2 == row["conj"], (row, s)


# ------------------------------------------------------------- permutations
def test_permutations_match_reference_fixtures(oracle):
    rows = _load("permutations.json.gz")["rows"]
    assert len(rows) > 5000
    for row in rows:
        got = oracle.permutations(row["hyper"], row["tl"], row["tr"], row["lrank"], row["rrank"], row["dig"],
                                  row["conj"])
        exp = {k: row[k] for k in ("ierr", "ncd", "nlu", "nru", "nhu")}
        assert {k: got[k] for k in exp} == exp, row
        if row["ierr"] == 0:
            assert got["lprm"][1:] == row["lprm"][1:], row
            assert got["rprm"][1:] == row["rprm"][1:], row
            assert got["dprm"][1:] == row["dprm"][1:], row


def _check_matricization(dig, lrank, rrank, perms):
    """Invariant: after the O2N permutations, L = (contracted, left-free) and
    R = (contracted, right-free) with matching contracted order, and D's N2O
    lists left-free then right-free positions."""
    lp, rp, dp = perms["lprm"], perms["rprm"], perms["dprm"]
    ncd = perms["ncd"]
    lcon = {lp[i + 1]: -dig[i] for i in range(lrank) if dig[i] < 0}
    rcon = {rp[j + 1]: j + 1 for j in range(rrank) if dig[lrank + j] < 0}
    assert sorted(lcon) == list(range(1, ncd + 1))
    assert sorted(rcon) == list(range(1, ncd + 1))
    for k in range(1, ncd + 1):
        assert lcon[k] == rcon[k]  # same contracted index at the same slot
    lfree = sorted((lp[i + 1], dig[i]) for i in range(lrank) if dig[i] > 0)
    rfree = sorted((rp[j + 1], dig[lrank + j]) for j in range(rrank) if dig[lrank + j] > 0)
    assert [d for _, d in lfree] + [d for _, d in rfree] == dp[1:]


@pytest.mark.parametrize("seed", range(4))
def test_permutations_matricize(oracle, seed):
    rows = _load("patterns.json")["rows"]
    rng = random.Random(seed)
    for row in rng.sample(rows, 200):
        if row["ierr"] or row["rrank"] <= 0 or row["lrank"] <= 0:
            continue
        perms = oracle.permutations(1, 1, 0, row["lrank"], row["rrank"], row["dig"], 0)
        if perms["ierr"] or perms["nhu"]:
            continue
        _check_matricization(row["dig"], row["lrank"], row["rrank"], perms)


# ---------------------------------------------------------------- tensor ops
@pytest.fixture(scope="module")
def ops():
    meta = _load("tensor_ops.json")["cases"]
    arrs = np.load(os.path.join(GOLD, "tensor_ops.npz"))
    return meta, arrs


def _is_conj_r_complex(case, oracle):
    e, dig, dr, lr, rr, cj = oracle.str2dig(case["pattern"])
    lconj, rconj = (cj >> 1) & 1, (cj >> 2) & 1
    if cj & 1:
        lconj, rconj = 1 - lconj, 1 - rconj
    return case["kind"] in (C4, C8) and rconj and lr > 0 and rr > 0


def test_copy_matches_reference_bit_exact(oracle, ops):
    meta, arrs = ops
    n = 0
    for c in meta:
        if c["op"] != "copy":
            continue
        assert c["ierr"] == 0
        out = oracle.copy(c["pattern"], c["ldims"], arrs[c["key"] + "_l"])
        assert np.array_equal(out.view(np.uint8), arrs[c["key"] + "_d"].view(np.uint8)), c
        n += 1
    assert n >= 40


def test_add_matches_reference(oracle, ops):
    meta, arrs = ops
    for c in meta:
        if c["op"] != "add":
            continue
        d = arrs[c["key"] + "_d0"].copy()
        s = complex(*c["scale"])
        oracle.add(c["pattern"], c["ldims"], d, arrs[c["key"] + "_l"], s)
        exp = arrs[c["key"] + "_d"]
        tol = 2e-7 if c["kind"] in (R4, C4) else 1e-15
        assert np.max(np.abs(d - exp)) <= tol * (1 + np.max(np.abs(exp))), c


def test_init_scale_match_reference(oracle, ops):
    meta, arrs = ops
    import ctypes as C
    for c in meta:
        if c["op"] != "init_scale":
            continue
        v = complex(*c["value"])
        d = arrs[c["key"] + "_d0"].copy()
        oracle.lib.cptal_init(c["kind"], d.size, d.ctypes.data, v.real, v.imag)
        assert np.array_equal(d, arrs[c["key"] + "_init"])
        d = arrs[c["key"] + "_d0"].copy()
        oracle.lib.cptal_scale(c["kind"], d.size, d.ctypes.data, v.real, v.imag)
        tol = 2e-7 if c["kind"] in (R4, C4) else 1e-15
        assert np.max(np.abs(d - arrs[c["key"] + "_scale"])) <= tol * 4, c


def _rel(a, b):
    nb = np.linalg.norm(b.astype(np.complex128))
    return np.linalg.norm(a.astype(np.complex128) - b.astype(np.complex128)) / max(nb, 1e-300)


def test_contract_matches_reference(oracle, ops):
    """Relative Frobenius error vs the reference's own output: <=1e-12 (R8/C8),
    <=1e-5 (R4/C4). Complex contractions with a conjugated right operand are
    excluded here: the reference's -DNO_BLAS build mis-orders that operand (its
    in-house GEMM assumes R(c,r) but the permutation planner produced R(r,c));
    those cases are pinned by the long-double direct evaluation instead."""
    meta, arrs = ops
    n = nconj = 0
    for c in meta:
        if c["op"] != "contract":
            continue
        k = c["key"]
        d = arrs[k + "_d0"].copy()
        e = oracle.contract(c["pattern"], d, c["ddims"], arrs[k + "_l"], c["ldims"], arrs[k + "_r"], c["rdims"],
                            complex(*c["alpha"]) if c["kind"] in (C4, C8) else c["alpha"][0], c["accumulative"])
        assert e == c["ierr"], c
        if e:
            continue
        exact = oracle.contract_exact(c["pattern"], arrs[k + "_d0"], c["ddims"], arrs[k + "_l"], c["ldims"],
                                      arrs[k + "_r"], c["rdims"],
                                      complex(*c["alpha"]) if c["kind"] in (C4, C8) else c["alpha"][0],
                                      c["accumulative"])
        tol = 1e-12 if c["kind"] in (R8, C8) else 1e-5
        assert _rel(d, exact) <= tol, c
        if _is_conj_r_complex(c, oracle):
            nconj += 1
            continue
        assert _rel(d, arrs[k + "_d"]) <= tol, c
        n += 1
    assert n > 100


def test_closed_form_test_cpp(oracle):
    """TALSH/test.cpp:92-147: D(a,b,i,j)+=L(c,b,d,a)*R(j,d,i,c)*2 with L=0.001,
    R=0.01 constant; every D element is 2*0.001*0.01*(c*d extent product)."""
    V = 4
    ddims = [V, V, V, V]
    l = np.full(V ** 4, 0.001); r = np.full(V ** 4, 0.01); d = np.zeros(V ** 4)
    assert oracle.contract("D(a,b,i,j)+=L(c,b,d,a)*R(j,d,i,c)", d, ddims, l, [V] * 4, r, [V] * 4, 2.0) == 0
    np.testing.assert_allclose(d, 2 * 0.001 * 0.01 * V * V, rtol=1e-14)
    # accumulate again on top
    assert oracle.contract("D(a,b,i,j)+=L(c,b,d,a)*R(j,d,i,c)", d, ddims, l, [V] * 4, r, [V] * 4, 2.0) == 0
    np.testing.assert_allclose(d, 2 * 2 * 0.001 * 0.01 * V * V, rtol=1e-14)


def test_copy_brute_force(oracle):
    rng = np.random.default_rng(3)
    for rank in range(1, 7):
        dims = [int(x) for x in rng.integers(1, 5, size=rank)]
        perm = rng.permutation(rank)
        l = rng.standard_normal(int(np.prod(dims)))
        out = oracle.permute(dims, [int(p) + 1 for p in perm], l)
        # numpy in C order: reverse dims; input axis k (DLF) goes to output axis perm[k]
        lt = l.reshape(dims[::-1])
        odims = [0] * rank
        for k in range(rank):
            odims[perm[k]] = dims[k]
        ref = np.empty(odims[::-1])
        # out index tuple (DLF) o: o[perm[k]] = i[k]
        it = np.nditer(lt, flags=["multi_index"])
        for x in it:
            i = it.multi_index[::-1]
            o = [0] * rank
            for k in range(rank):
                o[perm[k]] = i[k]
            ref[tuple(o[::-1])] = x
        assert np.array_equal(out, ref.ravel())


def test_contract_brute_force_einsum(oracle):
    """Small random contractions against numpy.einsum in float64."""
    rng = np.random.default_rng(11)
    pats = ["D(a,b,c,d)+=L(d,i,b,j)*R(j,c,i,a)", "D(a,b,i,j)+=L(a,k,i,l)*R(l,b,k,j)",
            "D(a,b)+=L(i,a)*R(b,i)", "D(x,y,z)+=L(z,q,x)*R(y,q)"]
    for p in pats:
        e, dig, dr, lr, rr, cj = oracle.str2dig(p)
        ext = {}
        ll = p[p.index("L(") + 2:p.index(")", p.index("L("))].split(",")
        rl = p[p.index("R(") + 2:p.index(")", p.index("R("))].split(",")
        dl = p[p.index("D(") + 2:p.index(")")].split(",")
        for x in set(ll + rl + dl):
            ext[x] = int(rng.integers(2, 6))
        ld = [ext[x] for x in ll]; rd = [ext[x] for x in rl]; dd = [ext[x] for x in dl]
        L = rng.standard_normal(int(np.prod(ld))); R = rng.standard_normal(int(np.prod(rd)))
        D = rng.standard_normal(int(np.prod(dd)))
        d = D.copy()
        assert oracle.contract(p, d, dd, L, ld, R, rd, 0.75) == 0
        # DLF arrays are numpy Fortran order
        Lf = L.reshape(ld, order="F"); Rf = R.reshape(rd, order="F"); Df = D.reshape(dd, order="F")
        sub = "".join(ll) + "," + "".join(rl) + "->" + "".join(dl)
        ref = Df + 0.75 * np.einsum(sub, Lf, Rf)
        np.testing.assert_allclose(d, ref.ravel(order="F"), rtol=1e-12, atol=1e-12)


def test_planted_mistake_is_caught(oracle):
    """A plausible mistake (transposed contracted operand order) must fail the
    einsum comparison: contract with the R indices swapped and check it differs."""
    rng = np.random.default_rng(5)
    L = rng.standard_normal(12); R = rng.standard_normal(12); D = np.zeros(9)
    good = D.copy(); bad = D.copy()
    oracle.contract("D(a,b)+=L(a,i)*R(i,b)", good, [3, 3], L, [3, 4], R, [4, 3])
    oracle.contract("D(a,b)+=L(a,i)*R(b,i)", bad, [3, 3], L, [3, 4], R.reshape(4, 3).T.ravel().copy() * 0 + R,
                    [3, 4])
    assert not np.allclose(good, bad)


# ------------------------------------------------------------ live reference
def test_random_patterns_vs_live_reference(oracle, reference):
    from golden.gen_golden import random_pattern
    rng = random.Random(99)
    for i in range(300):
        p = random_pattern(rng, hyper=(i % 3 == 0))
        if i % 5 == 0:
            k = rng.randrange(len(p)); p = p[:k] + p[k + 1:]
        assert oracle.str2dig(p) == reference.str2dig(p), p


def test_random_contractions_vs_live_reference(oracle, reference):
    from golden.gen_golden import random_pattern
    rng = random.Random(1234)
    nrs = np.random.default_rng(1234)
    done = 0
    while done < 40:
        p = random_pattern(rng, conj=False)
        e, dig, dr, lr, rr, cj = oracle.str2dig(p)
        if e or rr < 0:
            continue
        ext = {}
        ldims = []
        for k in range(lr):
            v = dig[k]
            key = ("d", v) if v > 0 else ("c", k)
            ext.setdefault(key, int(nrs.integers(1, 5)))
            ldims.append(ext[key])
        rdims = []
        for k in range(rr):
            v = dig[lr + k]
            key = ("d", v) if v > 0 else ("c", -v - 1)
            ext.setdefault(key, int(nrs.integers(1, 5)))
            rdims.append(ext[key])
        ddims = [ext[("d", q + 1)] for q in range(dr)]
        kind = (R8, C8)[done % 2]
        dt = NP_DTYPE[kind]
        mk = lambda n: (nrs.uniform(-1, 1, n) + (1j * nrs.uniform(-1, 1, n) if kind == C8 else 0)).astype(dt)
        D = mk(int(np.prod(ddims)) if dr else 1)
        L = mk(int(np.prod(ldims)) if lr else 1)
        R = mk(int(np.prod(rdims)) if rr else 1)
        d = D.copy()
        e1 = oracle.contract(p, d, ddims, L, ldims, R, rdims, 1.25)
        e2, out = reference.contract(p, D, ddims, L, ldims, R, rdims, 1.25)
        assert e1 == e2, p
        if e1 == 0:
            assert _rel(d, out) <= 1e-12, p
        done += 1


def _einsum_conj(pattern, D, dd, L, ld, R, rd, alpha, accumulative=True):
    """numpy restatement of the complex contraction with conjugation flags:
    L+ / R+ conjugate that operand; D+ conjugates both operands and not alpha
    (TALSH/tensor_algebra_cpu.F90:3827-3838). Computed in complex128 from
    the given (possibly complex64) values; DLF bodies are Fortran order."""
    import re
    m = re.fullmatch(r"D(\+?)\(([^)]*)\)\+=L(\+?)\(([^)]*)\)\*R(\+?)\(([^)]*)\)", pattern.replace(" ", ""))
    dc, dl, lc, ll, rc, rl = m.groups()
    lab = lambda s: s.split(",") if s else []
    dl, ll, rl = lab(dl), lab(ll), lab(rl)
    conj_l = bool(lc) != bool(dc)
    conj_r = bool(rc) != bool(dc)
    Lf = L.astype(np.complex128).reshape(ld, order="F")
    Rf = R.astype(np.complex128).reshape(rd, order="F")
    if conj_l:
        Lf = np.conj(Lf)
    if conj_r:
        Rf = np.conj(Rf)
    letters = {x: chr(ord("a") + i) for i, x in enumerate(dict.fromkeys(dl + ll + rl))}
    sub = "".join(letters[x] for x in ll) + "," + "".join(letters[x] for x in rl) + "->" + \
        "".join(letters[x] for x in dl)
    prod = alpha * np.einsum(sub, Lf, Rf)
    base = D.astype(np.complex128).reshape(dd, order="F") if accumulative else 0
    return (base + prod).ravel(order="F")


@pytest.mark.parametrize("kind", [C8, C4])
def test_conjugated_complex_contractions_vs_einsum(oracle, kind):
    """The conjugated complex forms, where the reference's -DNO_BLAS build
    cannot pin the oracle (R+ operands): the oracle's contraction and its
    long-double evaluator against numpy.einsum with np.conj, for R+, L+R+,
    D+ (which flips both operand flags) and mixed forms, accumulating and
    not, complex alpha. 1e-12 (C8), 1e-5 (C4, computed in complex64 by the
    oracle) relative Frobenius error."""
    rng = np.random.default_rng(77 + kind)
    pats = ["D(a,b)+=L(a,i)*R+(i,b)", "D(a,b)+=L+(a,i)*R+(i,b)", "D+(a,b)+=L(i,a)*R(b,i)",
            "D+(a,b)+=L(a,i)*R+(i,b)", "D+(a,b,c)+=L+(c,i,a)*R(b,i)", "D(a,b,c,d)+=L(d,i,b,j)*R+(j,c,i,a)",
            "D+(a,b,c,d)+=L+(d,i,b,j)*R+(j,c,i,a)", "D(a,b,i,j)+=L+(a,k,i,l)*R(l,b,k,j)",
            "D(x,y)+=L(x,q,p)*R+(p,y,q)", "D(a,b)+=L(a)*R+(b)"]
    dt = NP_DTYPE[kind]
    for p in pats:
        e, dig, dr, lr, rr, cj = oracle.str2dig(p)
        assert e == 0, p
        import re
        m = re.fullmatch(r"D\+?\(([^)]*)\)\+=L\+?\(([^)]*)\)\*R\+?\(([^)]*)\)", p)
        dl, ll, rl = (g.split(",") for g in m.groups())
        ext = {x: int(rng.integers(2, 7)) for x in set(dl + ll + rl)}
        dd = [ext[x] for x in dl]; ld = [ext[x] for x in ll]; rd = [ext[x] for x in rl]
        mk = lambda n: (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(dt)
        D, L, R = mk(int(np.prod(dd))), mk(int(np.prod(ld))), mk(int(np.prod(rd)))
        alpha = complex(0.6, -0.35)
        tol = 1e-12 if kind == C8 else 1e-5
        for acc in (True, False):
            want = _einsum_conj(p, D, dd, L, ld, R, rd, alpha, acc)
            d = D.copy()
            assert oracle.contract(p, d, dd, L, ld, R, rd, alpha, acc) == 0
            assert _rel(d, want) <= tol, (p, acc)
            ex = oracle.contract_exact(p, D, dd, L, ld, R, rd, alpha, acc)
            assert _rel(ex, want) <= 1e-12, (p, acc)


def test_einsum_conj_restatement_detects_a_wrong_flag(oracle):
    """The einsum restatement is sensitive to each flag: dropping R's
    conjugation, or conjugating alpha under D+, changes the result well
    beyond the tolerance."""
    rng = np.random.default_rng(3)
    mk = lambda n: rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    L, R, D = mk(12), mk(12), mk(9)
    good = _einsum_conj("D(a,b)+=L(a,i)*R+(i,b)", D, [3, 3], L, [3, 4], R, [4, 3], 0.5 + 0.5j)
    bad = _einsum_conj("D(a,b)+=L(a,i)*R(i,b)", D, [3, 3], L, [3, 4], R, [4, 3], 0.5 + 0.5j)
    assert _rel(bad, good) > 1e-3
    d = D.copy()
    assert oracle.contract("D+(a,b)+=L(a,i)*R(i,b)", d, [3, 3], L, [3, 4], R, [4, 3], 0.5 + 0.5j) == 0
    conj_alpha = _einsum_conj("D+(a,b)+=L(a,i)*R(i,b)", D, [3, 3], L, [3, 4], R, [4, 3], 0.5 - 0.5j)
    assert _rel(d, conj_alpha) > 1e-3
    assert _rel(d, _einsum_conj("D+(a,b)+=L(a,i)*R(i,b)", D, [3, 3], L, [3, 4], R, [4, 3], 0.5 + 0.5j)) <= 1e-12

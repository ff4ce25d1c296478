"""Benchmark of the MI355X TAL-SH tensor-block engine (one GPU).

Headline workload (one "step"): the configured DLF transposes through
talshTensorCopy with operands resident in HBM: rank-6 tensors with every
extent 24 over 64 seeded non-identity permutations and rank-4 tensors with
every extent 128 over all 23 non-identity permutations, each in R4 and R8.
Algorithmic bytes per transpose = 2 * volume * element size (read + write).

Also reported (not the headline value): dense contractions through
talshTensorContract (R8 on FP64 matrix cores, R4 native FP32, R4 three-way
BF16 split, C8) on D(a,b,i,j)+=L(a,k,i,l)*R(l,b,k,j) with a=b=i=j=96, k=l=128,
and D(a,b,c,d)+=L(d,i,b,j)*R(j,c,i,a) with extents 32; a batch of 4096
block contractions D(a,b,c,d)+=L(a,k,c,l)*R(l,b,k,d) (one grouped launch vs
one asynchronous talshTensorContract task per block); the learned plan
selector's regret against exhaustive timing on fresh contractions.

Prints ONE JSON line.
"""
import argparse
import ctypes as C
import json
import os
import random
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK = 78.6
FP32_PEAK = 157.3
BF16_PEAK = 2500.0
METRIC = "contraction TFLOP/s & %MFMA peak; transpose GB/s & %HBM; block-sparse blocks/s"


def perms_rank6(n=64, seed=2026):
    rng = random.Random(seed)
    out, seen = [], set()
    while len(out) < n:
        p = list(range(6))
        rng.shuffle(p)
        if p == sorted(p) or tuple(p) in seen:
            continue
        seen.add(tuple(p))
        out.append(p)
    return out


def perms_rank4():
    import itertools
    return [list(p) for p in itertools.permutations(range(4)) if list(p) != [0, 1, 2, 3]]


def copy_pattern(perm):
    n = len(perm)
    lab = [f"i{k}" for k in range(n)]
    dlab = [None] * n
    for k in range(n):
        dlab[perm[k]] = lab[k]
    return ("D(" + ",".join(dlab) + ")=L(" + ",".join(lab) + ")").encode()


class GpuBench:
    def __init__(self):
        from exatensor_amd import talsh as T
        self.T = T
        self.eng = T.Engine(gpus=[0], arena_bytes=int(os.environ.get("BENCH_ARENA_BYTES", 64 << 30)))
        self.L = self.eng.lib
        self.gpu = self.eng.gpu(0)

    def tensor(self, kind, dims, init=0.0):
        t = self.T.Tensor(self.L, kind, dims, self.gpu, init)
        if t.err:
            raise RuntimeError(f"talshTensorConstruct failed: {t.err}")
        return t

    def fill(self, t, seed):
        """Seeded U[-1,1] body generated on the host, uploaded once."""
        from exatensor_amd import talsh as T
        rng = np.random.default_rng(seed)
        a = rng.uniform(-1, 1, t.volume).astype(T.NP_DTYPE[t.kind]) if t.kind in (T.R4, T.R8) else \
            (rng.standard_normal(t.volume) + 1j * rng.standard_normal(t.volume)).astype(T.NP_DTYPE[t.kind])
        t.upload(a)
        return a

    def sync(self):
        self.L.talshAmdDeviceSynchronize(0)

    def prof(self, cls):
        ms, n = C.c_double(), C.c_longlong()
        self.L.talshAmdProfileQuery(cls, C.byref(ms), C.byref(n))
        return ms.value, n.value


def run_transposes(gb, steps, warmup):
    T = gb.T
    work = []  # (label, kind, dims, perms)
    for kind in (T.R4, T.R8):
        work.append(("rank6_24", kind, [24] * 6, perms_rank6()))
        work.append(("rank4_128", kind, [128] * 4, perms_rank4()))
    tens = {}
    for label, kind, dims, _ in work:
        src = gb.tensor(kind, dims)
        gb.fill(src, 7 if kind == T.R8 else 5)
        dst = gb.tensor(kind, dims)
        tens[(label, kind)] = (src, dst)
    gb.sync()
    pats = {(label, kind): [copy_pattern(p) for p in perms] for label, kind, dims, perms in work}

    def one_step():
        nbytes = 0
        for label, kind, dims, perms in work:
            src, dst = tens[(label, kind)]
            es = 4 if kind == T.R4 else 8
            for pat in pats[(label, kind)]:
                e = gb.L.talshTensorCopy(pat, dst.ptr, src.ptr, gb.gpu, T.DEV_NULL, T.COPY_MT, None)
                if e:
                    raise RuntimeError(f"talshTensorCopy {pat} -> {e}")
                nbytes += 2 * src.volume * es
        return nbytes

    for _ in range(warmup):
        one_step()
    gb.sync()
    gb.L.talshAmdProfile(1)
    gb.prof(T.KERNEL_TRANSPOSE)  # drop warmup records
    t0 = time.perf_counter()
    total = 0
    for _ in range(steps):
        total += one_step()
    gb.sync()
    t1 = time.perf_counter()
    kms, nl = gb.prof(T.KERNEL_TRANSPOSE)
    gb.L.talshAmdProfile(0)
    for src, dst in tens.values():
        src.destruct(); dst.destruct()
    per_launch_bytes = total / max(nl, 1)
    return dict(bytes=total, seconds=t1 - t0, kernel_ms=kms, launches=nl,
                achieved_gbs=(total / 1e9) / (kms / 1e3) if kms > 0 else None,
                per_launch_bytes=per_launch_bytes, transposes_per_step=sum(len(w[3]) for w in work))


def copy_probe(gb, reps=3):
    """Identity copies of a 2 GiB R8 tensor (k_copy16: reads and writes
    exactly 2 GiB each). Used only to calibrate hardware byte counters:
    the HBM bytes a profiler's FETCH_SIZE/WRITE_SIZE report for this kernel
    are known, so their ratio scales the counters of the transpose kernel."""
    T = gb.T
    src = gb.tensor(T.R8, [1 << 14, 1 << 14], 1.0)
    dst = gb.tensor(T.R8, [1 << 14, 1 << 14], 0.0)
    for _ in range(reps):
        e = gb.L.talshTensorCopy(b"D(a,b)=L(a,b)", dst.ptr, src.ptr, gb.gpu, T.DEV_NULL, T.COPY_MT, None)
        if e:
            raise RuntimeError(f"identity copy -> {e}")
    gb.sync()
    src.destruct(); dst.destruct()


def measured_traffic():
    """HBM bytes per k_permute_tiled launch over the transpose workload, as
    measured by hardware counters and recorded in
    profiles/k_permute_tiled_transposes_traffic.md (None if absent)."""
    p = os.path.join(ROOT, "profiles", "k_permute_tiled_transposes_traffic.md")
    try:
        with open(p) as f:
            for ln in f:
                if ln.startswith("traffic_bytes_per_launch:"):
                    return int(float(ln.split(":", 1)[1]))
    except OSError:
        pass
    return None


def sampled_error_c3(l, r, d, n=256, seed=5):
    """Relative error of D(a,b,i,j) = sum_{k,l} L(a,k,i,l) * R(l,b,k,j)
    (a=b=i=j=96, k=l=128, DLF bodies) at n seeded positions against float64
    (complex128) sums of the same inputs computed here with numpy."""
    Lf = l.reshape(96, 128, 96, 128, order="F")
    Rf = r.reshape(128, 96, 128, 96, order="F")
    Df = d.reshape(96, 96, 96, 96, order="F")
    idx = np.random.default_rng(seed).integers(0, 96, (n, 4))
    ct = np.complex128 if np.iscomplexobj(l) else np.float64
    want = np.array([np.sum(Lf[a, :, i, :].astype(ct) * Rf[:, b, :, j].astype(ct).T) for a, b, i, j in idx])
    got = Df[idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3]].astype(ct)
    return float(np.linalg.norm(got - want) / np.linalg.norm(want))


def run_contractions(gb, reps=3):
    T = gb.T
    out = {}
    cases = [("c3_r8", T.R8, False), ("c3_r4_fp32", T.R4, False), ("c3_r4_bf16x3", T.R4, True), ("c3_c8", T.C8, False),
             ("c3_c4_fp32", T.C4, False), ("c3_c4_bf16x3", T.C4, True)]
    pat3 = b"D(a,b,i,j)+=L(a,k,i,l)*R(l,b,k,j)"
    for name, kind, fast in cases:
        d = gb.tensor(kind, [96] * 4)
        l = gb.tensor(kind, [96, 128, 96, 128]); r = gb.tensor(kind, [128, 96, 128, 96])
        lh = gb.fill(l, 1); rh = gb.fill(r, 2)
        cplx = kind in (T.C4, T.C8)
        flops = (8.0 if cplx else 2.0) * 96 ** 4 * 128 ** 2
        if fast:
            gb.L.talshEnableFastMath(T.DEV_AMD_GPU, 0)
        gb.sync()
        e = gb.L.talshTensorContract(pat3, d.ptr, l.ptr, r.ptr, 1.0, 0.0, gb.gpu, T.DEV_NULL, T.COPY_MTT, T.NOPE, None)
        if e:
            raise RuntimeError(f"contract {name}: {e}")
        gb.L.talshAmdProfile(1)
        gb.prof(T.KERNEL_GEMM); gb.prof(T.KERNEL_TRANSPOSE)
        t0 = time.perf_counter()
        for _ in range(reps):
            gb.L.talshTensorContract(pat3, d.ptr, l.ptr, r.ptr, 1.0, 0.0, gb.gpu, T.DEV_NULL, T.COPY_MTT, T.NOPE, None)
        gb.sync()
        t1 = time.perf_counter()
        gms, gn = gb.prof(T.KERNEL_GEMM)
        tms, tn = gb.prof(T.KERNEL_TRANSPOSE)
        gb.L.talshAmdProfile(0)
        gb.L.talshDisableFastMath(T.DEV_AMD_GPU, 0)
        peak = FP64_PEAK if kind in (T.R8, T.C8) else (BF16_PEAK if fast else FP32_PEAK)
        tf = flops * reps / (t1 - t0) / 1e12
        gemm_tf = flops * gn / (gms / 1e3) / 1e12 if gms > 0 else None
        # outside the timed region: the result of the last (non-accumulating) call against float64 sums
        err = sampled_error_c3(lh, rh, d.download())
        progress(f"contraction {name}: {tf:.1f} TFLOP/s")
        out[name] = dict(tflops=round(tf, 3), gemm_tflops=round(gemm_tf, 3) if gemm_tf else None,
                         gemm_frac_peak=round(gemm_tf / peak, 4) if gemm_tf else None, peak_tflops=peak,
                         ms_per_op=round((t1 - t0) / reps * 1e3, 3), gemm_ms=round(gms / max(gn, 1), 3),
                         transpose_ms=round(tms / max(gn, 1), 3), rel_err_fp64_sampled=float("%.3g" % err))
        del lh, rh
        for t in (d, l, r):
            t.destruct()
    # configured CPU-runnable case, extents 32
    pat1 = b"D(a,b,c,d)+=L(d,i,b,j)*R(j,c,i,a)"
    d = gb.tensor(T.R8, [32] * 4); l = gb.tensor(T.R8, [32] * 4); r = gb.tensor(T.R8, [32] * 4)
    gb.fill(l, 3); gb.fill(r, 4)
    n1 = 20

    def time_c1():
        gb.L.talshTensorContract(pat1, d.ptr, l.ptr, r.ptr, 1.0, 0.0, gb.gpu, T.DEV_NULL, T.COPY_MTT, T.YEP, None)
        gb.sync()
        t0 = time.perf_counter()
        for _ in range(n1):
            gb.L.talshTensorContract(pat1, d.ptr, l.ptr, r.ptr, 1.0, 0.0, gb.gpu, T.DEV_NULL, T.COPY_MTT, T.YEP, None)
        gb.sync()
        return (time.perf_counter() - t0) / n1

    dt = time_c1()
    v, t = C.c_int(), C.c_int()
    gb.L.talshAmdSelectorChoice(pat1, 4, T.ints([32] * 4), 4, T.ints([32] * 4), 4, T.ints([32] * 4), T.R8, T.YEP, 0,
                                C.byref(v), C.byref(t))
    # the same calls planned by the transpose-only byte-count heuristic
    mode = gb.L.talshAmdGetPlanMode()
    gb.L.talshAmdSetPlanMode(0)
    try:
        dt0 = time_c1()
    finally:
        gb.L.talshAmdSetPlanMode(mode)
    out["c1_r8_ext32"] = dict(tflops=round(2.0 * 32 ** 6 / dt / 1e12, 3), ms_per_op=round(dt * 1e3, 3),
                              plan=dict(variant=v.value, tile=t.value, fused=v.value >= T.FUSED_VARIANT),
                              heuristic_ms_per_op=round(dt0 * 1e3, 3))
    for t in (d, l, r):
        t.destruct()
    return out


def _finish(gb, tk):
    st = C.c_int()
    e = gb.L.talshTaskWait(C.byref(tk), C.byref(st))
    if e or st.value != gb.T.TALSH_TASK_COMPLETED:
        raise RuntimeError(f"task wait: {e} status {st.value}")
    gb.L.talshTaskDestruct(C.byref(tk))


def run_blocksparse(gb, reps=3, per_block_leg=True):
    """C4: ~4096 uneven block contractions of D(a,b,c,d)+=L(a,k,c,l)*R(l,b,k,d)
    (7 segments per mode, extents from {8..64}, 20% block fill), as one
    persistent grouped launch replayed from a hipGraph, against one
    talshTensorContract per block."""
    from exatensor_amd import blocksparse as BS
    T = gb.T
    ext, ops = BS.generate()
    out = {"ops": len(ops), "flops": BS.flops(ext, ops)}
    for kind, fast, name in ((T.R8, 0, "r8"), (T.R4, 0, "r4_fp32"), (T.R4, 1, "r4_bf16x3")):
        if fast:  # R4 three-way BF16 split for the batch and the per-block baseline alike
            gb.L.talshEnableFastMath(T.DEV_AMD_GPU, 0)
        tens = {}

        def get(tag, key, modes, init):
            if (tag, key) not in tens:
                tens[(tag, key)] = gb.tensor(kind, BS.dims(ext, key, modes), init)
            return tens[(tag, key)]
        D = [get("D", o[0], "abcd", 0.0) for o in ops]
        L = [get("L", o[1], "akcl", 0.01) for o in ops]
        R = [get("R", o[2], "lbkd", 0.02) for o in ops]
        gb.sync()
        n = len(ops)
        arr = lambda ts: (T._TP * n)(*[C.pointer(t.t) for t in ts])
        b = C.c_void_p()
        e = gb.L.talshAmdBatchCreate(BS.PATTERN.encode(), n, arr(D), arr(L), arr(R), 1.0, T.YEP, 0, C.byref(b))
        if e:
            raise RuntimeError(f"talshAmdBatchCreate: {e}")
        gb.L.talshAmdBatchExecute(b, 1)
        gb.L.talshAmdProfile(1)
        gb.prof(T.KERNEL_GROUPED)
        t0 = time.perf_counter()
        for _ in range(reps):
            gb.L.talshAmdBatchExecute(b, 1)
        t1 = time.perf_counter()
        kms, kn = gb.prof(T.KERNEL_GROUPED)
        gb.L.talshAmdProfile(0)
        gb.L.talshAmdBatchDestroy(b)
        grouped_s = (t1 - t0) / reps
        # baseline: one asynchronous talshTensorContract task (its own
        # transposes + GEMM) per block, up to 64 tasks in flight over the
        # engine's streams, as a block-level worker issues them
        pat = BS.PATTERN.encode()

        def per_block(count):
            window, inflight = 64, []
            for i in range(count):
                tk = T.TalshTask()
                gb.L.talshTaskClean(C.byref(tk))
                while True:
                    e = gb.L.talshTensorContract(pat, D[i].ptr, L[i].ptr, R[i].ptr, 1.0, 0.0, gb.gpu, T.DEV_NULL,
                                                 T.COPY_MTT, T.YEP, C.byref(tk))
                    # an operand still held by an earlier task (blocks chained on one D): retire the oldest
                    # task in flight and issue again
                    if e in (T.TALSH_IN_PROGRESS, T.TRY_LATER) and inflight:
                        _finish(gb, inflight.pop(0))
                        continue
                    break
                if e:
                    raise RuntimeError(f"per-block contract {i}: {e}")
                inflight.append(tk)
                if len(inflight) >= window:
                    _finish(gb, inflight.pop(0))
            for tk in inflight:
                _finish(gb, tk)

        out[name] = dict(grouped_blocks_per_s=round(n / grouped_s, 1), grouped_ms=round(grouped_s * 1e3, 3),
                         grouped_kernel_ms=round(kms / max(kn, 1), 3),
                         grouped_tflops=round(out["flops"] / grouped_s / 1e12, 3))
        if per_block_leg:
            # an untimed pass first: it fills the plan memos (the autotuner plans
            # block shapes outside the learned selector's domain on first sight)
            per_block(n)
            gb.sync()
            t0 = time.perf_counter()
            per_block(n)
            gb.sync()
            per_block_s = time.perf_counter() - t0
            out[name].update(per_block_blocks_per_s=round(n / per_block_s, 1),
                             per_block_ms=round(per_block_s * 1e3, 3), speedup=round(per_block_s / grouped_s, 2))
        if fast:
            gb.L.talshDisableFastMath(T.DEV_AMD_GPU, 0)
        progress(f"block-sparse {name}: grouped {grouped_s * 1e3:.1f} ms"
                 + (f", per block {per_block_s * 1e3:.1f} ms" if per_block_leg else ""))
        for t in tens.values():
            t.destruct()
    return out


def run_selector(gb, n=300, seed=303):
    """C5: regret of the committed plan-selector model on n fresh seeded
    contractions (disjoint seed from training), against exhaustive timing of
    every candidate, next to the byte-count heuristic's regret."""
    from exatensor_amd import selector as S
    if not os.path.exists(S.DEFAULT_MODEL):
        return {"skipped": "no trained model"}
    rng = np.random.default_rng(seed)
    specs, times = [], []
    t0 = time.perf_counter()
    for i in range(n):
        spec = S.random_contraction(rng)
        _, _, _, tm = S.time_candidates(gb.L, gb.gpu, spec)
        specs.append(spec)
        times.append(tm)
        if (i + 1) % 50 == 0:
            progress(f"plan selector: {i + 1}/{n} contractions timed")
    m = S.Mlp(gb.L, path=S.DEFAULT_MODEL)
    rep = S.regret_report(gb.L, specs, times, m)
    gb.L.talshAmdSelectorInstall(None)
    m.close()
    rep["seed"] = seed
    rep["timing_seconds"] = round(time.perf_counter() - t0, 1)
    return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in rep.items()}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(class_seconds=4.0, contract_seconds=3.0):
    """The reference TAL-SH CPU path (oracle/_ref, built from the reference's
    own sources) timed on this host's cores for a bounded sample of the same
    transposes: each of the four classes (rank-6 extent-24 and rank-4
    extent-128, R8 and R4) gets about class_seconds, at least one
    permutation; falls back to the oracle's C restatement if the reference
    build is absent. With the reference build present, also times the
    extent-32 rank-4 R8 contraction on DEV_HOST (returned second, None
    otherwise)."""
    ncores = os.cpu_count() or 1
    os.environ.setdefault("OMP_NUM_THREADS", str(ncores))
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libtalsh_ref.so")
    classes = [("rank6_24_r8", 8, [24] * 6, perms_rank6()), ("rank6_24_r4", 4, [24] * 6, perms_rank6()),
               ("rank4_128_r8", 8, [128] * 4, perms_rank4()), ("rank4_128_r4", 4, [128] * 4, perms_rank4())]

    class Tens(C.Structure):
        _fields_ = [("a", C.c_void_p), ("b", C.c_void_p), ("c", C.c_void_p), ("d", C.c_void_p),
                    ("e", C.c_int), ("f", C.c_int)]
    kind = "reference" if os.path.exists(ref_so) else "port"
    if kind == "reference":
        L = C.CDLL(ref_so)
        sz = C.c_size_t(1 << 24); mx = C.c_int()
        L.talshInit(C.byref(sz), C.byref(mx), 0, None, 0, None, 0, None)
        ierr = C.c_int()
        L.talshSetMemAllocPolicyHost(0, 1, C.byref(ierr))
        L.talshTensorConstruct.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                           C.c_void_p, C.c_double, C.c_double]
        L.talshTensorCopy.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.talshTensorGetBodyAccess.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int]
    else:
        lib = C.CDLL(os.path.join(ROOT, "oracle", "libcptal.so"))
    p = C.c_void_p()
    total_bytes, total_s, per_class, parts = 0, 0.0, {}, []
    for label, es, dims, perms in classes:
        rank = len(dims)
        vol = int(np.prod(dims))
        dk = 8 if es == 8 else 4  # TAL-SH data kinds R8 = 8, R4 = 4
        x = np.random.default_rng(5).uniform(-1, 1, vol).astype(np.float64 if es == 8 else np.float32)
        dimsc = (C.c_int * rank)(*dims)
        if kind == "reference":
            src, dst = Tens(), Tens()
            for t in (src, dst):
                L.talshTensorClean(C.byref(t))
                assert L.talshTensorConstruct(C.byref(t), dk, rank, dimsc, 0, None, -1, None, 0.0, 0.0) == 0
            L.talshTensorGetBodyAccess(C.byref(src), C.byref(p), dk, 0, 0)
            C.memmove(p, x.ctypes.data, x.nbytes)

            def one(perm):
                return L.talshTensorCopy(copy_pattern(perm), C.byref(dst), C.byref(src), 0, 0, 6, None)
        else:
            out = np.empty_like(x)

            def one(perm):
                o2n = (C.c_int * (rank + 1))(1, *[q + 1 for q in perm])
                return lib.cptal_copy_dlf(dk, rank, dimsc, o2n, C.c_void_p(x.ctypes.data),
                                          C.c_void_p(out.ctypes.data), 0)
        one(perms[0])  # warm-up
        t0 = time.perf_counter()
        n = 0
        while n < len(perms) and (n == 0 or time.perf_counter() - t0 < class_seconds):
            e = one(perms[n])
            if e:
                raise RuntimeError(f"CPU baseline copy failed: {e}")
            n += 1
        dt = time.perf_counter() - t0
        nbytes = 2 * vol * es * n
        total_bytes += nbytes
        total_s += dt
        per_class[label] = round(nbytes / 1e9 / dt, 3)
        parts.append(f"{n}/{len(perms)} {label}")
        if kind == "reference":
            for t in (src, dst):
                L.talshTensorDestruct(C.byref(t))
        del x
    cores = int(os.environ["OMP_NUM_THREADS"])
    base = dict(value=round(total_bytes / 1e9 / total_s, 3), unit="GB/s", cores=cores, kind=kind,
                cpu_model=cpu_model(), per_class_gbs=per_class,
                sample="permutations via talshTensorCopy on DEV_HOST, about %.0f s per class: %s (%.1f s)"
                       % (class_seconds, ", ".join(parts), total_s))
    c1 = None
    if kind == "reference":
        # D(a,b,c,d)+=L(d,i,b,j)*R(j,c,i,a), extents 32, on the reference's CPU path
        L.talshTensorContract.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                          C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
        d4 = (C.c_int * 4)(32, 32, 32, 32)
        ct = [Tens() for _ in range(3)]
        for t in ct:
            L.talshTensorClean(C.byref(t))
            assert L.talshTensorConstruct(C.byref(t), 8, 4, d4, 0, None, -1, None, 0.0, 0.0) == 0
        rng = np.random.default_rng(34)
        for t in ct[1:]:
            v = rng.uniform(-1, 1, 32 ** 4)
            L.talshTensorGetBodyAccess(C.byref(t), C.byref(p), 8, 0, 0)
            C.memmove(p, v.ctypes.data, v.nbytes)
        pat1 = b"D(a,b,c,d)+=L(d,i,b,j)*R(j,c,i,a)"

        def contract():
            e = L.talshTensorContract(pat1, C.byref(ct[0]), C.byref(ct[1]), C.byref(ct[2]), 1.0, 0.0, 0, 0, 26, 1,
                                      None)
            if e:
                raise RuntimeError(f"CPU baseline contraction failed: {e}")
        contract()
        t0 = time.perf_counter()
        reps = 0
        while reps < 20 and (reps < 1 or time.perf_counter() - t0 < contract_seconds):
            contract()
            reps += 1
        dt = time.perf_counter() - t0
        c1 = dict(gflops=round(2.0 * 32 ** 6 * reps / dt / 1e9, 3), ms_per_op=round(dt / reps * 1e3, 3),
                  cores=int(os.environ["OMP_NUM_THREADS"]), kind="reference",
                  sample=f"{reps} blocking talshTensorContract calls on DEV_HOST ({dt:.1f} s)")
        for t in ct:
            L.talshTensorDestruct(C.byref(t))
    return base, c1


def progress(msg):
    """One line per finished part on stderr (the JSON result stays alone on stdout)."""
    print(f"bench: {msg} ({time.strftime('%H:%M:%S')})", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-contract", action="store_true", help="skip the contraction components")
    ap.add_argument("--no-blocksparse", action="store_true", help="skip the block-sparse component")
    ap.add_argument("--no-selector", action="store_true", help="skip the plan-selector regret component")
    ap.add_argument("--no-per-block", action="store_true",
                    help="block-sparse component without its one-task-per-block baseline leg (for counter runs)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the timed region, run identity copies of known size (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this engine runs one process's share of tensor blocks on one MI355X; "
              "multi-GPU sharding is a design note, not a measured configuration.", file=sys.stderr)
        sys.exit(2)
    gb = GpuBench()
    tr = run_transposes(gb, args.steps, args.warmup)
    progress("transposes timed")
    if args.copy_probe:
        copy_probe(gb)
    comps = {} if args.no_contract else run_contractions(gb)
    progress("contractions timed")
    if not args.no_blocksparse:
        comps["blocksparse"] = run_blocksparse(gb, per_block_leg=not args.no_per_block)
        progress("block-sparse batches timed")
    if not args.no_selector:
        comps["plan_selector"] = run_selector(gb)
        progress("plan selector timed")
    value = tr["bytes"] / 1e9 / tr["seconds"]
    achieved = tr["achieved_gbs"]
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tr["seconds"] / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded U[-1,1])",
        "config": {"workload": "DLF transposes: rank-6 extent-24 x 64 seeded non-identity permutations and "
                               "rank-4 extent-128 x all 23 non-identity permutations, R4 and R8, via talshTensorCopy",
                   "transposes_per_step": tr["transposes_per_step"], "bytes_per_step": tr["bytes"] // args.steps},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": measured_traffic(), "kernel": "k_permute_tiled", "launches": tr["launches"],
                     "avg_launch_ms": round(tr["kernel_ms"] / max(tr["launches"], 1), 4)},
        "components": comps,
    }
    if not args.no_cpu:
        line["cpu_baseline"], c1_cpu = cpu_baseline()
        if c1_cpu and "c1_r8_ext32" in comps:
            comps["c1_r8_ext32"]["cpu_reference"] = c1_cpu
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

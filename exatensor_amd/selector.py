"""Learned contraction-plan selector: workload generator, timing collection,
GPU training and regret evaluation.

The engine's planner enumerates every distinct matricization of a GEMM-shaped
contraction (contracted order from L or R, free orders from the operand or
from D, D or D^T) and every fused-gather plan (no transposes: the product
reads the tensors through offset tables), times every GEMM tile shape of the
data kind. The selector is a small MLP that predicts log(kernel time) of each
candidate from its features (talshAmdPlanFeatures) and picks the minimum, so
it chooses the operand order, fused vs. separate transposes and the tile. It replaces fixed rules
of the kind the reference hard-codes (TALSH/talshc.cpp:366-488,
TALSH/tensor_algebra_gpu_nvidia.cu:1188-1272, TALSH/tensor_algebra_cpu.F90:9291-9298).

Labels are measured by the engine itself: every candidate of a contraction
is executed through talshTensorContract with talshAmdSetPlanVariant, and the
GPU time of its transpose and GEMM kernels is read from the engine's
per-launch HIP events (talshAmdProfileQuery). That exhaustive search is also
the yardstick of the selector's regret:
    regret = t(selected) / t(best candidate) - 1.

Command line (runs on a GPU):
    python -m exatensor_amd.selector collect --n 2000 --seed 11 --out train.npz
    python -m exatensor_amd.selector train --data train.npz --model plan_model.txt
    python -m exatensor_amd.selector evaluate --data heldout.npz --model plan_model.txt
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

from . import talsh as T

LETTERS = "abcdefghijklmnopqrstuvwxyz"
KINDS = (T.R8, T.R4, T.C8)
KIND_P = (0.45, 0.45, 0.10)
DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "plan_model.txt")


def random_contraction(rng, log2_vol=(16, 26), max_log2_flops=35.0):
    """A seeded random GEMM-shaped contraction.

    Returns dict(kind, acc, pattern, ddims, ldims, rdims): tensor ranks 2..6,
    at least one contracted and one free index per operand, extents
    log-uniform in [2, 96], the largest operand's volume log-uniform in
    2^log2_vol, at most 2^max_log2_flops real flops.
    """
    kind = int(rng.choice(KINDS, p=KIND_P))
    acc = bool(rng.random() < 0.7)
    while True:
        nc, nl, nr = (int(v) for v in rng.integers(1, 4, size=3))
        if nc + nl > 6 or nc + nr > 6 or nl + nr > 6 or nl + nr < 2:
            continue
        n = nc + nl + nr
        target = rng.uniform(*log2_vol)
        ext = np.rint(np.exp(rng.uniform(math.log(2), math.log(96), n))).astype(int)
        lab = list(LETTERS[:n])
        cl, ll, rl = lab[:nc], lab[nc:nc + nl], lab[nc + nl:]
        e = dict(zip(lab, ext.tolist()))
        vol = lambda ls: float(np.prod([e[x] for x in ls]))
        vmax = max(vol(cl + ll), vol(cl + rl), vol(ll + rl))
        if abs(math.log2(vmax) - target) > 0.5:
            continue
        if math.log2(2.0 * vol(lab)) > max_log2_flops:
            continue
        L = list(rng.permutation(cl + ll))
        R = list(rng.permutation(cl + rl))
        D = list(rng.permutation(ll + rl))
        pat = f"D({','.join(D)})+=L({','.join(L)})*R({','.join(R)})"
        return dict(kind=kind, acc=acc, pattern=pat, ddims=[e[x] for x in D], ldims=[e[x] for x in L],
                    rdims=[e[x] for x in R])


def plan_features(lib, spec, fast_math=0):
    """Candidates of a contraction: (variants, tiles, features[ncand][nf])."""
    nf = lib.talshAmdPlanFeatureCount()
    cap = 256
    v = (C.c_int * cap)()
    t = (C.c_int * cap)()
    f = (C.c_float * (cap * nf))()
    n = C.c_int()
    e = lib.talshAmdPlanFeatures(spec["pattern"].encode(), len(spec["ddims"]), T.ints(spec["ddims"]),
                                 len(spec["ldims"]), T.ints(spec["ldims"]), len(spec["rdims"]), T.ints(spec["rdims"]),
                                 spec["kind"], T.YEP if spec["acc"] else T.NOPE, fast_math, cap, v, t, f, C.byref(n))
    if e:
        raise RuntimeError(f"talshAmdPlanFeatures {spec['pattern']}: {e}")
    k = min(n.value, cap)
    feats = np.ctypeslib.as_array(f)[:k * nf].reshape(k, nf).copy()
    return np.array(v[:k], dtype=np.int32), np.array(t[:k], dtype=np.int32), feats


def _kernel_ms(lib):
    tot = 0.0
    for cls in (T.KERNEL_TRANSPOSE, T.KERNEL_GEMM):
        ms, n = C.c_double(), C.c_longlong()
        lib.talshAmdProfileQuery(cls, C.byref(ms), C.byref(n))
        tot += ms.value
    return tot


def time_candidates(lib, gpu, spec, reps=2):
    """Exhaustive timing: GPU kernel time (ms, min over reps after one warm-up
    run) of every candidate of one contraction."""
    variants, tiles, feats = plan_features(lib, spec)
    kind = spec["kind"]
    D = T.Tensor(lib, kind, spec["ddims"], gpu, 0.0)
    L = T.Tensor(lib, kind, spec["ldims"], gpu, 0.5)
    R = T.Tensor(lib, kind, spec["rdims"], gpu, 0.25)
    for x in (D, L, R):
        if x.err:
            raise RuntimeError(f"construct: {x.err}")
    pat = spec["pattern"].encode()
    acc = T.YEP if spec["acc"] else T.NOPE
    times = np.empty(len(variants))
    lib.talshAmdProfile(1)
    try:
        for i, (v, t) in enumerate(zip(variants, tiles)):
            lib.talshAmdSetPlanVariant(int(v), int(t))
            best = float("inf")
            for r in range(reps + 1):
                _kernel_ms(lib)
                e = lib.talshTensorContract(pat, D.ptr, L.ptr, R.ptr, 1.0, 0.0, gpu, T.DEV_NULL, T.COPY_MTT, acc, None)
                if e:
                    raise RuntimeError(f"contract {spec['pattern']} variant {v} tile {t}: {e}")
                ms = _kernel_ms(lib)
                if r > 0:
                    best = min(best, ms)
            times[i] = best
    finally:
        lib.talshAmdSetPlanVariant(-1, -1)
        lib.talshAmdProfile(0)
        for x in (D, L, R):
            x.destruct()
    return variants, tiles, feats, times


def collect(n, seed, reps=2, log=None):
    """Time every candidate of n seeded random contractions on GPU 0."""
    eng = T.Engine(gpus=[0])
    lib, gpu = eng.lib, eng.gpu(0)
    rng = np.random.default_rng(seed)
    specs, times, t0 = [], [], time.time()
    for i in range(n):
        spec = random_contraction(rng)
        _, _, _, tm = time_candidates(lib, gpu, spec, reps)
        specs.append(spec)
        times.append(tm)
        if log and (i + 1) % 100 == 0:
            print(f"collect: {i + 1}/{n} contractions, {time.time() - t0:.0f} s", file=log, flush=True)
    return specs, times


def save_dataset(path, specs, times):
    np.savez_compressed(path, specs=np.array(json.dumps(specs)),
                        times=np.array(json.dumps([t.tolist() for t in times])))


def load_dataset(path):
    z = np.load(path, allow_pickle=False)
    specs = json.loads(str(z["specs"]))
    times = [np.array(t) for t in json.loads(str(z["times"]))]
    return specs, times


def design_matrix(lib, specs, times, clip=None):
    """Stacked features and log-time labels of every candidate; group offsets.
    clip: labels of candidates slower than clip x the contraction's fastest
    are capped there (the model need not rank hopeless candidates precisely)."""
    X, y, groups = [], [], [0]
    for spec, tm in zip(specs, times):
        _, _, f = plan_features(lib, spec)
        if len(f) != len(tm):
            raise RuntimeError(f"candidate count changed for {spec['pattern']}")
        X.append(f)
        t = np.maximum(tm, 1e-4)
        if clip:
            t = np.minimum(t, clip * t.min())
        y.append(np.log(t))
        groups.append(groups[-1] + len(tm))
    return np.concatenate(X).astype(np.float32), np.concatenate(y).astype(np.float32), groups


class Mlp:
    """Handle of an engine-side MLP (talshAmdMlp*)."""

    def __init__(self, lib, nin=None, hidden=2, width=64, seed=1, path=None):
        self.lib = lib
        self.h = C.c_void_p()
        if path:
            e = lib.talshAmdMlpLoad(path.encode(), C.byref(self.h))
        else:
            e = lib.talshAmdMlpCreate(nin, hidden, width, seed, C.byref(self.h))
        if e:
            raise RuntimeError(f"MLP create/load failed: {e}")
        n = C.c_int()
        lib.talshAmdMlpNumParams(self.h, C.byref(n))
        self.nparams = n.value

    def params(self):
        p = np.empty(self.nparams, np.float32)
        self.lib.talshAmdMlpGetParams(self.h, p.ctypes.data_as(T._FP))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self.lib.talshAmdMlpSetParams(self.h, p.ctypes.data_as(T._FP))

    def train(self, X, y, epochs=60, batch=512, lr=2e-3, seed=7):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        loss = np.empty(epochs, np.float32)
        e = self.lib.talshAmdMlpTrain(self.h, X.ctypes.data_as(T._FP), y.ctypes.data_as(T._FP), len(y), epochs, batch,
                                      lr, seed, loss.ctypes.data_as(T._FP))
        if e:
            raise RuntimeError(f"talshAmdMlpTrain: {e}")
        return loss

    def predict(self, X):
        X = np.ascontiguousarray(X, np.float32)
        out = np.empty(len(X), np.float32)
        e = self.lib.talshAmdMlpPredict(self.h, X.ctypes.data_as(T._FP), len(X), out.ctypes.data_as(T._FP))
        if e:
            raise RuntimeError(f"talshAmdMlpPredict: {e}")
        return out

    def gradient(self, X, y):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.empty(self.nparams, np.float32)
        loss = C.c_float()
        e = self.lib.talshAmdMlpGradient(self.h, X.ctypes.data_as(T._FP), y.ctypes.data_as(T._FP), len(y),
                                         g.ctypes.data_as(T._FP), C.byref(loss))
        if e:
            raise RuntimeError(f"talshAmdMlpGradient: {e}")
        return g, loss.value

    def save(self, path):
        if self.lib.talshAmdMlpSave(self.h, path.encode()):
            raise RuntimeError(f"cannot write {path}")

    def install(self):
        return self.lib.talshAmdSelectorInstall(self.h)

    def close(self):
        if self.h:
            self.lib.talshAmdMlpDestroy(self.h)
            self.h = C.c_void_p()


def choice(lib, spec):
    """(variant, tile) that talshTensorContract would run for this contraction."""
    v, t = C.c_int(), C.c_int()
    e = lib.talshAmdSelectorChoice(spec["pattern"].encode(), len(spec["ddims"]), T.ints(spec["ddims"]),
                                   len(spec["ldims"]), T.ints(spec["ldims"]), len(spec["rdims"]), T.ints(spec["rdims"]),
                                   spec["kind"], T.YEP if spec["acc"] else T.NOPE, 0, C.byref(v), C.byref(t))
    if e:
        raise RuntimeError(f"talshAmdSelectorChoice: {e}")
    return v.value, max(t.value, 0)


def regret_report(lib, specs, times, model):
    """Regret of the model's choices and of the byte-count heuristic's
    (variant 0 with its default tile), against exhaustive search. Leaves the
    model installed."""
    lib.talshAmdSelectorInstall(None)
    heur_pick = [choice(lib, s) for s in specs]
    model.install()
    sel, heur, tsel, theur, tbest = [], [], 0.0, 0.0, 0.0
    for spec, tm, hp in zip(specs, times, heur_pick):
        variants, tiles, _ = plan_features(lib, spec)
        mp = choice(lib, spec)
        index = {(int(a), int(b)): i for i, (a, b) in enumerate(zip(variants, tiles))}
        best = float(np.min(tm))
        ts, th = tm[index[mp]], tm[index[hp]]
        sel.append(ts / best - 1.0)
        heur.append(th / best - 1.0)
        tsel += ts
        theur += th
        tbest += best
    sel, heur = np.array(sel), np.array(heur)
    return dict(contractions=len(specs), mean_regret=float(sel.mean()), median_regret=float(np.median(sel)),
                p90_regret=float(np.percentile(sel, 90)), total_time_regret=tsel / tbest - 1.0,
                heuristic_mean_regret=float(heur.mean()), heuristic_total_time_regret=theur / tbest - 1.0,
                optimal_picks=float(np.mean(sel <= 1e-9)))


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m exatensor_amd.selector")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--n", type=int, required=True)
    c.add_argument("--seed", type=int, required=True)
    c.add_argument("--reps", type=int, default=2)
    c.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("--data", nargs="+", required=True)
    t.add_argument("--model", required=True)
    t.add_argument("--hidden", type=int, default=2)
    t.add_argument("--width", type=int, default=64)
    t.add_argument("--epochs", type=int, default=60)
    t.add_argument("--batch", type=int, default=512)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--clip", type=float, default=None)
    ev = sub.add_parser("evaluate")
    ev.add_argument("--data", nargs="+", required=True)
    ev.add_argument("--model", default=DEFAULT_MODEL)
    a = ap.parse_args(argv)
    if a.cmd == "collect":
        specs, times = collect(a.n, a.seed, a.reps, log=sys.stderr)
        save_dataset(a.out, specs, times)
        print(json.dumps({"collected": len(specs), "candidates": int(sum(len(x) for x in times))}))
        return 0
    eng = T.Engine(gpus=[0])
    lib = eng.lib
    specs, times = [], []
    for p in a.data:
        s_, t_ = load_dataset(p)
        specs += s_
        times += t_
    if a.cmd == "train":
        X, y, _ = design_matrix(lib, specs, times, a.clip)
        m = Mlp(lib, X.shape[1], a.hidden, a.width)
        t0 = time.time()
        loss = m.train(X, y, a.epochs, a.batch, a.lr)
        m.save(a.model)
        print(json.dumps({"rows": len(y), "contractions": len(specs), "epochs": a.epochs,
                          "train_seconds": round(time.time() - t0, 2), "first_loss": float(loss[0]),
                          "final_loss": float(loss[-1]), "train_regret": regret_report(lib, specs, times, m)}))
        m.close()
        return 0
    m = Mlp(lib, path=a.model)
    print(json.dumps(regret_report(lib, specs, times, m)))
    m.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

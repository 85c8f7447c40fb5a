"""Learned surrogate for the GA: an ensemble of MLPs (nin -> 256 -> 256 -> 256 ->
nout, ReLU) that maps a normalised conductance vector (log10 of the free
parameters, scaled to [-1, 1]) to the GA's score features.  Training and
inference run on libneurogpu's BF16 MFMA kernels with FP32 accumulation.

The surrogate only pre-screens offspring: ``screen`` ranks candidates by the
ensemble-mean predicted score and keeps the top fraction plus every
candidate whose ensemble disagreement is high, and those are simulated; the
GA's scores always come from the simulator."""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from .ga import GA_FEATURES, PENALTY, Target

_FP = C.POINTER(C.c_float)


def _bind():
    L = _lib.lib()
    if getattr(L, "_mlp_bound", False):
        return L
    vp = C.c_void_p
    L.ngpu_mlp_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]
    L.ngpu_mlp_create.restype = C.c_int
    L.ngpu_mlp_nparams.argtypes = [vp]
    L.ngpu_mlp_nparams.restype = C.c_longlong
    for n in ("ngpu_mlp_set_weights", "ngpu_mlp_get_weights"):
        getattr(L, n).argtypes = [vp, _FP]
        getattr(L, n).restype = C.c_int
    L.ngpu_mlp_train.argtypes = [vp, _FP, _FP, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint, _FP]
    L.ngpu_mlp_train.restype = C.c_int
    L.ngpu_mlp_predict.argtypes = [vp, _FP, C.c_int, _FP]
    L.ngpu_mlp_predict.restype = C.c_int
    L.ngpu_mlp_gradients.argtypes = [vp, C.c_int, _FP, _FP, C.c_int, _FP, _FP]
    L.ngpu_mlp_gradients.restype = C.c_int
    L.ngpu_mlp_last_error.argtypes = [vp]
    L.ngpu_mlp_last_error.restype = C.c_char_p
    L.ngpu_mlp_close.argtypes = [vp]
    L.ngpu_mlp_close.restype = None
    L.ngpu_gemm_bf16.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.POINTER(C.c_ushort), C.POINTER(C.c_ushort), _FP]
    L.ngpu_gemm_bf16.restype = C.c_int
    L._mlp_bound = True
    return L


def to_bf16_bits(x: np.ndarray) -> np.ndarray:
    """Round float32 to bfloat16 (nearest even), returned as uint16 bit patterns."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    nan = np.isnan(np.asarray(x, dtype=np.float32))
    r[nan] = 0x7FC0
    return r


def from_bf16_bits(b: np.ndarray) -> np.ndarray:
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32)


def gemm_bf16(A_bits: np.ndarray, B_bits: np.ndarray, ta: bool, tb: bool, M: int, N: int, K: int, device: int = 0):
    L = _bind()
    A = np.ascontiguousarray(A_bits, dtype=np.uint16)
    B = np.ascontiguousarray(B_bits, dtype=np.uint16)
    C_ = np.empty((M, N), dtype=np.float32)
    rc = L.ngpu_gemm_bf16(device, int(ta), int(tb), M, N, K, A.ctypes.data_as(C.POINTER(C.c_ushort)),
                          B.ctypes.data_as(C.POINTER(C.c_ushort)), C_.ctypes.data_as(_FP))
    if rc != 0:
        raise _lib.NgpuError(rc, "gemm failed")
    return C_


class MLPEnsemble:
    def __init__(self, nin: int, nout: int, hidden: int = 256, members: int = 5, seed: int = 0, device: int = 0):
        self.L = _bind()
        self.nin, self.nout, self.H, self.members = nin, nout, hidden, members
        self.h = C.c_void_p()
        rc = self.L.ngpu_mlp_create(device, nin, hidden, nout, members, C.byref(self.h))
        if rc != 0:
            raise _lib.NgpuError(rc, "ngpu_mlp_create failed")
        self.nparam = int(self.L.ngpu_mlp_nparams(self.h))
        self.set_weights(self.init_weights(seed))
        self.xmu = self.xsd = self.ymu = self.ysd = None

    def _check(self, rc):
        if rc != 0:
            raise _lib.NgpuError(rc, (self.L.ngpu_mlp_last_error(self.h) or b"").decode())

    def shapes(self):
        H, i, o = self.H, self.nin, self.nout
        return [(H, i), (H,), (H, H), (H,), (H, H), (H,), (o, H), (o,)]

    def init_weights(self, seed: int) -> np.ndarray:
        """He-uniform weights, zero biases, per member (seeded)."""
        rng = np.random.default_rng(seed)
        out = []
        for _ in range(self.members):
            for shp in self.shapes():
                if len(shp) == 2:
                    lim = np.sqrt(6.0 / shp[1])
                    out.append(rng.uniform(-lim, lim, size=shp).ravel())
                else:
                    out.append(np.zeros(shp))
        return np.concatenate(out).astype(np.float32)

    def set_weights(self, w: np.ndarray) -> None:
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.nparam * self.members
        self._check(self.L.ngpu_mlp_set_weights(self.h, w.ctypes.data_as(_FP)))

    def get_weights(self) -> np.ndarray:
        w = np.empty(self.nparam * self.members, dtype=np.float32)
        self._check(self.L.ngpu_mlp_get_weights(self.h, w.ctypes.data_as(_FP)))
        return w

    def unpack(self, w: np.ndarray, member: int):
        base = member * self.nparam
        out, o = [], base
        for shp in self.shapes():
            n = int(np.prod(shp))
            out.append(w[o:o + n].reshape(shp))
            o += n
        return out

    # -------------------------------------------------------- raw (standardised) I/O
    def gradients(self, member: int, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(self.nparam, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self._check(self.L.ngpu_mlp_gradients(self.h, member, x.ctypes.data_as(_FP), t.ctypes.data_as(_FP),
                                              x.shape[0], g.ctypes.data_as(_FP), loss.ctypes.data_as(_FP)))
        return g, float(loss[0])

    def train_raw(self, x, t, epochs=200, batch=512, lr=1e-3, seed=0):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        losses = np.zeros(self.members * epochs, dtype=np.float32)
        self._check(self.L.ngpu_mlp_train(self.h, x.ctypes.data_as(_FP), t.ctypes.data_as(_FP), x.shape[0], epochs,
                                          batch, lr, seed, losses.ctypes.data_as(_FP)))
        return losses.reshape(self.members, epochs)

    def predict_raw(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty((self.members, x.shape[0], self.nout), dtype=np.float32)
        self._check(self.L.ngpu_mlp_predict(self.h, x.ctypes.data_as(_FP), x.shape[0], out.ctypes.data_as(_FP)))
        return out

    # -------------------------------------------------------- normalised I/O
    def fit(self, x: np.ndarray, y: np.ndarray, **kw):
        """x: normalised genomes [n, nin]; y: target features [n, nout]
        (standardised internally)."""
        self.ymu = y.mean(axis=0)
        self.ysd = y.std(axis=0) + 1e-9
        return self.train_raw(x, (y - self.ymu) / self.ysd, **kw)

    def refit(self, x: np.ndarray, y: np.ndarray, **kw):
        """Continue training from the current weights with the scaling of the
        first fit (the screened GA's surrogate learning from new simulations)."""
        if self.ymu is None:
            return self.fit(x, y, **kw)
        return self.train_raw(x, (y - self.ymu) / self.ysd, **kw)

    def predict(self, x: np.ndarray):
        """Ensemble mean and standard deviation of the de-standardised outputs."""
        p = self.predict_raw(x).astype(np.float64) * self.ysd + self.ymu
        return p.mean(axis=0), p.std(axis=0)

    def close(self):
        if self.h and self.h.value:
            self.L.ngpu_mlp_close(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def genome_to_input(g: np.ndarray, lo: np.ndarray, hi: np.ndarray) -> np.ndarray:
    """log10 genomes in [lo, hi] -> [-1, 1]."""
    return (2.0 * (g - lo) / (hi - lo) - 1.0).astype(np.float32)


def feature_targets(filled: np.ndarray) -> np.ndarray:
    """Surrogate outputs: the GA features of every stimulus, flattened."""
    return filled[..., GA_FEATURES].reshape(filled.shape[0], -1)


def predicted_scores(ens: MLPEnsemble, x_in: np.ndarray, target: Target, nstim: int):
    mu, sd = ens.predict(x_in)
    mu = mu.reshape(-1, nstim, len(GA_FEATURES))
    sd = sd.reshape(-1, nstim, len(GA_FEATURES))
    scale = target.scale[list(GA_FEATURES)]
    d = np.abs(mu - target.features[None, :, GA_FEATURES]) / scale
    score = np.minimum(d, PENALTY).sum(axis=(1, 2))
    disagreement = (sd / scale).sum(axis=(1, 2))
    return score, disagreement


def make_screen(ens: MLPEnsemble, target: Target, lo, hi, nstim: int, keep_frac: float = 0.1,
                disagree_frac: float = 0.02, explore: float = 0.0, random_frac: float = 0.0, seed: int = 0):
    """Screen for run_ga: keep the best ``keep_frac`` by predicted score minus
    ``explore`` times the ensemble disagreement (an optimistic bound; 0 ranks by
    the ensemble mean) plus the ``disagree_frac`` most uncertain candidates
    (the disagreement gate: where the ensemble does not agree the candidate
    goes to full simulation) plus ``random_frac`` drawn uniformly from the rest
    (unscreened exploration)."""
    rng = np.random.default_rng(seed)

    def screen(genomes: np.ndarray) -> np.ndarray:
        s, d = predicted_scores(ens, genome_to_input(genomes, lo, hi), target, nstim)
        n = genomes.shape[0]
        k = max(1, int(round(keep_frac * n)))
        best = np.argsort(s - explore * d, kind="stable")[:k]
        kd = int(round(disagree_frac * n))
        unsure = np.argsort(-d, kind="stable")[:kd] if kd > 0 else np.array([], dtype=int)
        picked = np.unique(np.concatenate([best, unsure]))
        kr = int(round(random_frac * n))
        if kr > 0:
            rest = np.setdiff1d(np.arange(n), picked)
            picked = np.concatenate([picked, rng.choice(rest, size=min(kr, rest.size), replace=False)])
        return picked
    return screen


def bench(nin: int = 12, nout: int = 20, members: int = 5, ntrain: int = 65536, epochs: int = 10,
          npredict: int = 163840, seed: int = 0) -> dict:
    """Time the surrogate's two GPU workloads at the GA study's sizes: training
    the ensemble on ``ntrain`` samples (batch 512) and ranking ``npredict``
    candidates with every member.  Returns wall times and the GEMM FLOP
    rate of each (2 M N K per product, forward and backward)."""
    import time
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, size=(ntrain, nin)).astype(np.float32)
    t = np.tanh(x @ rng.normal(size=(nin, nout))).astype(np.float32)
    ens = MLPEnsemble(nin, nout, members=members, seed=seed)
    H = ens.H
    fwd = 2 * (nin * H + 2 * H * H + H * nout)             # flops per sample, forward
    ens.train_raw(x[:4096], t[:4096], epochs=1, batch=512, lr=1e-3, seed=1)   # warm-up
    t0 = time.perf_counter()
    ens.train_raw(x, t, epochs=epochs, batch=512, lr=1e-3, seed=2)
    train_s = time.perf_counter() - t0
    steps = epochs * (ntrain // 512)
    # a step: forward, then backward = two products per layer (the first layer's input gradient is skipped)
    train_flops = steps * 512 * members * (fwd + 2 * fwd - 2 * nin * H)
    xp = rng.uniform(-1, 1, size=(npredict, nin)).astype(np.float32)
    ens.predict_raw(xp[:1024])
    t0 = time.perf_counter()
    ens.predict_raw(xp)
    pred_s = time.perf_counter() - t0
    ens.close()
    return {"members": members, "train_samples": ntrain, "epochs": epochs, "train_s": train_s,
            "train_gemm_tflops": train_flops / train_s / 1e12, "predict_rows": npredict, "predict_s": pred_s,
            "predict_gemm_tflops": npredict * members * fwd / pred_s / 1e12}


if __name__ == "__main__":
    import json
    print(json.dumps(bench()), flush=True)

"""Learned warm start for constant fitting (no counterpart in the reference).

A small MLP predicts starting values for the LEARNABLE_CONSTANT slots of an
expression from its postfix opcodes and its targets, and the batched L-BFGS
(`optimize`, symbolic_torch/optimization.py:9-112 in the reference) starts
from the prediction instead of randn wherever that start has the lower loss.

Network (all layers `W (out, in)`, exact-erf GELU):
    x   = one-hot(Ops[:, :64] over opcodes 0..V-1)  ++  sign(y) log1p|y| at
          the first 256 target points  [++ (y - mean) / std at the same points
          (clipped to +-10) ++ sign(mean) log1p|mean|, log1p(std) when
          std_features]; zero padded to a multiple of 64.  The standardised
          block lowers the training loss but did not raise the fitted count
          on the benchmark (12674 vs 12605 of 16384) and widened the spread
          of fit times, so it is off by default.
    h_l = gelu(h_{l-1} W_l^T + b_l), l = 1..3, width 512
    out = h_3 W_4^T + b_4   (64 slots; only the constant slots are used)
with V = VAR_START_ID + n_x opcodes (53 for n_x = 3, the reference's
`n_operators`, src/pcfg.cpp:228).  Weights and activations are BF16 on the
MFMA units with FP32 accumulation; master weights and Adam moments are FP32.

The first layer never multiplies the one-hot zeros: the MFMA GEMM covers the
target columns only (K = 256), and the one-hot block adds, per row, the 64
weight rows it selects (from a transposed BF16 copy of W_1).  The weight
gradient of the first layer is still the dense dZ_1^T x product.

Every kernel is hand-written HIP in csrc/mlp_kernels.hip (GEMM on
v_mfma_f32_32x32x16_bf16, one-hot gather + GELU, transposes, GELU backward,
column sums, masked MSE head, Adam, feature build, prediction mask and start
selection); torch only allocates device memory and provides the stream.
"""
from __future__ import annotations

import math
import time
from typing import Dict, List, Optional, Tuple

import torch

from . import _lib
from ._C import Operator, _require_gpu
from .fused import fused_root

EPI_STORE, EPI_BIAS, EPI_BIAS_GELU = 0, 1, 2
N_OUT = 64  # output slots = maximum padded expression length the network reads


def _round_up(a: int, m: int) -> int:
    return (a + m - 1) // m * m


def _ck(code: int, what: str) -> None:
    _lib.check_status(int(code), what)


class WarmStartMLP:
    """The warm-start network, its Adam state and its activation buffers."""

    def __init__(self, device, n_x: int = 3, slots: int = 64, n_pts: int = 256, hidden: int = 512,
                 depth: int = 3, seed: int = 0, std_features: bool = False):
        if hidden % 32 or hidden <= 0 or depth < 1 or not 0 < slots <= N_OUT or n_pts < 0:
            raise ValueError("hidden must be a positive multiple of 32, depth >= 1, 0 < slots <= 64")
        self.device = torch.device(device)
        self.n_x, self.slots, self.n_pts = int(n_x), int(slots), int(n_pts)
        self.vocab = int(Operator.VAR_START_ID) + self.n_x
        self.std_block = 1 if std_features else 0
        self.row_len = _round_up(self.slots * self.vocab + self.n_pts * (1 + self.std_block) + 2 * self.std_block, 64)
        self.dims = [self.row_len] + [int(hidden)] * int(depth) + [N_OUT]
        gen = torch.Generator().manual_seed(int(seed))
        dev = self.device
        self.W: List[torch.Tensor] = []
        self.b: List[torch.Tensor] = []
        for i in range(len(self.dims) - 1):
            bound = 1.0 / math.sqrt(self.dims[i])
            self.W.append(((torch.rand(self.dims[i + 1], self.dims[i], generator=gen) * 2 - 1) * bound).to(dev))
            self.b.append(((torch.rand(self.dims[i + 1], generator=gen) * 2 - 1) * bound).to(dev))
        self.Wb = [torch.empty(w.shape, dtype=torch.bfloat16, device=dev) for w in self.W]
        # first layer's BF16 weights transposed (one row per input column): the
        # one-hot block of the input is applied as a sum of gathered rows
        self.WT0 = torch.empty(self.dims[0], self.dims[1], dtype=torch.bfloat16, device=dev)
        self.noop_suffix = torch.empty(self.slots + 1, self.dims[1], dtype=torch.float32, device=dev)
        self.onehot = self.slots * self.vocab
        self.mW = [torch.zeros_like(w) for w in self.W]
        self.vW = [torch.zeros_like(w) for w in self.W]
        self.mb = [torch.zeros_like(b) for b in self.b]
        self.vb = [torch.zeros_like(b) for b in self.b]
        self.gW = [torch.empty_like(w) for w in self.W]
        self.gb = [torch.empty_like(b) for b in self.b]
        self.step = 0
        self._bufs: Dict[int, Dict[str, object]] = {}
        self.refresh_bf16()

    # ------------------------------------------------------------------ utils
    @property
    def layers(self) -> int:
        return len(self.W)

    def _stream(self) -> int:
        return _lib.stream_of(self.device)

    def refresh_bf16(self) -> None:
        """Round the FP32 master weights into their BF16 copies."""
        lib = _lib.hip()
        for w, wb in zip(self.W, self.Wb):
            _ck(lib.sr_mlp_to_bf16(_lib.ptr(w), w.numel(), _lib.ptr(wb), self._stream()), "sr_mlp_to_bf16")
        self._refresh_wt0()

    def _refresh_wt0(self) -> None:
        _ck(_lib.hip().sr_mlp_transpose(_lib.ptr(self.Wb[0]), _lib.ptr(self.WT0), self.dims[1], self.dims[0],
                                        self._stream()), "sr_mlp_transpose")

    def _buffers(self, Bp: int, train: bool) -> Dict[str, object]:
        key = Bp * 2 + int(train)
        buf = self._bufs.get(key)
        if buf is not None:
            return buf
        dev, H = self.device, self.dims[1:-1]
        bf, f32 = torch.bfloat16, torch.float32
        buf = {
            "x": torch.empty(Bp, self.row_len, dtype=bf, device=dev),
            "z": [torch.empty(Bp, h, dtype=f32, device=dev) for h in H],
            "h": [torch.empty(Bp, h, dtype=bf, device=dev) for h in H],
            "out": torch.empty(Bp, N_OUT, dtype=f32, device=dev),
        }
        if train:
            wmax = max(self.dims)
            hmax = max(H)
            lib = _lib.hip()
            buf.update(
                dO=torch.empty(Bp, N_OUT, dtype=f32, device=dev),
                dOb=torch.empty(Bp, N_OUT, dtype=bf, device=dev),
                loss=torch.empty(Bp, dtype=f32, device=dev),
                hT=torch.empty(wmax * Bp, dtype=bf, device=dev),
                dzT=torch.empty(wmax * Bp, dtype=bf, device=dev),
                WT=torch.empty(max(self.dims[i] * self.dims[i + 1] for i in range(1, self.layers)) if self.layers > 1
                               else 1, dtype=bf, device=dev),
                dH=torch.empty(Bp * hmax, dtype=f32, device=dev),
                dZ=[torch.empty(Bp * hmax, dtype=f32, device=dev) for _ in range(2)],
                dZb=[torch.empty(Bp * hmax, dtype=bf, device=dev) for _ in range(2)],
                cs=torch.empty(max(int(lib.sr_mlp_col_sum_workspace_bytes(Bp, wmax)), 8), dtype=torch.uint8,
                               device=dev),
            )
        if len(self._bufs) > 4:
            self._bufs.clear()
        self._bufs[key] = buf
        return buf

    def _check_inputs(self, ops: torch.Tensor, Y: torch.Tensor) -> Tuple[int, int]:
        _require_gpu(ops, "Ops")
        if ops.dtype != torch.int64 or ops.dim() != 2 or not ops.is_contiguous():
            raise RuntimeError("Ops must be a contiguous (B, M) int64 tensor")
        if Y.dtype != torch.float32 or Y.dim() != 2 or not Y.is_contiguous() or Y.shape[0] != ops.shape[0]:
            raise RuntimeError("Y must be a contiguous (B, N) float32 tensor")
        if Y.shape[1] < self.n_pts:
            raise RuntimeError(f"the network reads the first {self.n_pts} target points; Y has {Y.shape[1]}")
        B, M = ops.shape
        if M > N_OUT:
            raise RuntimeError(f"the network predicts at most {N_OUT} constant slots (padded length {M})")
        return B, M

    # ---------------------------------------------------------------- forward
    def _forward(self, ops: torch.Tensor, Y: torch.Tensor, train: bool):
        B, M = self._check_inputs(ops, Y)
        Bp = _round_up(max(B, 1), 32)
        buf = self._buffers(Bp, train)
        lib, s = _lib.hip(), self._stream()
        _ck(lib.sr_mlp_features(_lib.ptr(ops), M, B, Bp, self.slots, self.vocab, _lib.ptr(Y), Y.shape[1],
                                self.n_pts, self.std_block, self.row_len, _lib.ptr(buf["x"]), s), "sr_mlp_features")
        # layer 1: MFMA product over the target columns, then the one-hot
        # columns as gathered weight rows, bias and GELU
        x, H1, oh = buf["x"], self.dims[1], self.onehot
        ky = self.row_len - oh
        _ck(lib.sr_mlp_gemm_nt(x.data_ptr() + 2 * oh, self.Wb[0].data_ptr() + 2 * oh, None, Bp, H1, ky,
                               self.row_len, self.row_len, EPI_STORE, _lib.ptr(buf["z"][0]), None, None, s),
            "sr_mlp_gemm_nt")
        # the trailing NO_OP slots (program padding) as one precomputed sum
        _ck(lib.sr_mlp_noop_suffix(_lib.ptr(self.WT0), self.slots, self.vocab, M, H1, _lib.ptr(self.noop_suffix), s),
            "sr_mlp_noop_suffix")
        _ck(lib.sr_mlp_onehot_gelu(_lib.ptr(buf["z"][0]), _lib.ptr(self.WT0), _lib.ptr(self.b[0]), _lib.ptr(ops), M, B,
                                   Bp, self.slots, self.vocab, H1, _lib.ptr(self.noop_suffix), _lib.ptr(buf["z"][0]),
                                   _lib.ptr(buf["h"][0]), s),
            "sr_mlp_onehot_gelu")
        h = buf["h"][0]
        for l in range(1, self.layers):
            n_in, n_out = self.dims[l], self.dims[l + 1]
            if l < self.layers - 1:
                _ck(lib.sr_mlp_gemm_nt(_lib.ptr(h), _lib.ptr(self.Wb[l]), _lib.ptr(self.b[l]), Bp, n_out, n_in, n_in,
                                       n_in, EPI_BIAS_GELU, None, _lib.ptr(buf["h"][l]), _lib.ptr(buf["z"][l]), s),
                    "sr_mlp_gemm_nt")
                h = buf["h"][l]
            else:
                _ck(lib.sr_mlp_gemm_nt(_lib.ptr(h), _lib.ptr(self.Wb[l]), _lib.ptr(self.b[l]), Bp, n_out, n_in, n_in,
                                       n_in, EPI_BIAS, _lib.ptr(buf["out"]), None, None, s), "sr_mlp_gemm_nt")
        return B, M, Bp, buf

    def forward(self, ops: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
        """Raw network outputs (B, 64), FP32."""
        B, _, _, buf = self._forward(ops, Y, False)
        return buf["out"][:B]

    def predict(self, ops: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
        """Predicted constants (B, M): the outputs on the LEARNABLE_CONSTANT
        slots, 0 on every other slot."""
        B, M, _, buf = self._forward(ops, Y, False)
        C = torch.empty((B, M), dtype=torch.float32, device=self.device)
        _ck(_lib.hip().sr_mlp_predict(_lib.ptr(buf["out"]), _lib.ptr(ops), M, B, _lib.ptr(C), self._stream()),
            "sr_mlp_predict")
        return C

    # --------------------------------------------------------------- training
    def loss_and_grads(self, ops: torch.Tensor, Y: torch.Tensor, label: torch.Tensor) -> torch.Tensor:
        """Forward + backward on one batch.  Fills self.gW / self.gb with the
        gradient of  sum_b sum_{k: Ops[b,k]==1} (out[b,k] - label[b,k])^2 / B
        and returns the per-row loss terms (B,) on the device."""
        B, M, Bp, buf = self._forward(ops, Y, True)
        if label.shape != (B, M) or label.dtype != torch.float32 or not label.is_contiguous():
            raise RuntimeError("label must be a contiguous (B, M) float32 tensor")
        lib, s = _lib.hip(), self._stream()
        _ck(lib.sr_mlp_mse_head(_lib.ptr(buf["out"]), _lib.ptr(label), _lib.ptr(ops), M, B, Bp, 1.0 / max(B, 1),
                                _lib.ptr(buf["dO"]), _lib.ptr(buf["dOb"]), _lib.ptr(buf["loss"]), s),
            "sr_mlp_mse_head")
        dz_f, dz_b = buf["dO"], buf["dOb"]
        cs = buf["cs"]
        for l in range(self.layers - 1, -1, -1):
            n_in, n_out = self.dims[l], self.dims[l + 1]
            h = buf["x"] if l == 0 else buf["h"][l - 1]
            _ck(lib.sr_mlp_col_sum(_lib.ptr(dz_f), Bp, n_out, _lib.ptr(self.gb[l]), _lib.ptr(cs), cs.numel(), s),
                "sr_mlp_col_sum")
            _ck(lib.sr_mlp_transpose(_lib.ptr(dz_b), _lib.ptr(buf["dzT"]), Bp, n_out, s), "sr_mlp_transpose")
            _ck(lib.sr_mlp_transpose(_lib.ptr(h), _lib.ptr(buf["hT"]), Bp, n_in, s), "sr_mlp_transpose")
            _ck(lib.sr_mlp_gemm_nt(_lib.ptr(buf["dzT"]), _lib.ptr(buf["hT"]), None, n_out, n_in, Bp, Bp, Bp, EPI_STORE,
                                   _lib.ptr(self.gW[l]), None, None, s), "sr_mlp_gemm_nt")
            if l == 0:
                break
            _ck(lib.sr_mlp_transpose(_lib.ptr(self.Wb[l]), _lib.ptr(buf["WT"]), n_out, n_in, s), "sr_mlp_transpose")
            _ck(lib.sr_mlp_gemm_nt(_lib.ptr(dz_b), _lib.ptr(buf["WT"]), None, Bp, n_in, n_out, n_out, n_out, EPI_STORE,
                                   _lib.ptr(buf["dH"]), None, None, s), "sr_mlp_gemm_nt")
            nxt = l % 2
            _ck(lib.sr_mlp_gelu_bwd(_lib.ptr(buf["dH"]), _lib.ptr(buf["z"][l - 1]), Bp * n_in,
                                    _lib.ptr(buf["dZ"][nxt]), _lib.ptr(buf["dZb"][nxt]), s), "sr_mlp_gelu_bwd")
            dz_f, dz_b = buf["dZ"][nxt], buf["dZb"][nxt]
        return buf["loss"][:B]

    def loss_rows(self, ops: torch.Tensor, Y: torch.Tensor, label: torch.Tensor) -> torch.Tensor:
        """Per-row loss terms sum_{k: Ops[b,k]==1} (out[b,k] - label[b,k])^2
        (forward only)."""
        B, M, Bp, buf = self._forward(ops, Y, False)
        rows = buf.get("loss_eval")
        if rows is None or rows.numel() < Bp:
            rows = buf["loss_eval"] = torch.empty(Bp, dtype=torch.float32, device=self.device)
        _ck(_lib.hip().sr_mlp_mse_head(_lib.ptr(buf["out"]), _lib.ptr(label), _lib.ptr(ops), M, B, Bp, 1.0,
                                       None, None, _lib.ptr(rows), self._stream()), "sr_mlp_mse_head")
        return rows[:B]

    def adam_step(self, lr: float, betas=(0.9, 0.999), eps: float = 1e-8, weight_decay: float = 0.0) -> None:
        """Apply the gradients in self.gW / self.gb (one Adam step) and refresh
        the BF16 weights."""
        self.step += 1
        lib, s = _lib.hip(), self._stream()
        for l in range(self.layers):
            _ck(lib.sr_mlp_adam(_lib.ptr(self.W[l]), _lib.ptr(self.gW[l]), _lib.ptr(self.mW[l]),
                                _lib.ptr(self.vW[l]), self.W[l].numel(), lr, betas[0], betas[1], eps, self.step,
                                weight_decay, _lib.ptr(self.Wb[l]), s), "sr_mlp_adam")
            _ck(lib.sr_mlp_adam(_lib.ptr(self.b[l]), _lib.ptr(self.gb[l]), _lib.ptr(self.mb[l]),
                                _lib.ptr(self.vb[l]), self.b[l].numel(), lr, betas[0], betas[1], eps, self.step,
                                weight_decay, None, s), "sr_mlp_adam")
        self._refresh_wt0()

    def train_step(self, ops, Y, label, lr: float, weight_decay: float = 0.0) -> torch.Tensor:
        rows = self.loss_and_grads(ops, Y, label)
        self.adam_step(lr, weight_decay=weight_decay)
        return rows

    # ------------------------------------------------------------- persistence
    def state_dict(self) -> Dict[str, torch.Tensor]:
        d = {"meta": torch.tensor([self.n_x, self.slots, self.n_pts, self.dims[1], self.layers - 1, self.step,
                                   self.std_block])}
        for l in range(self.layers):
            d[f"W{l}"] = self.W[l].cpu()
            d[f"b{l}"] = self.b[l].cpu()
        return d

    def load_state_dict(self, d: Dict[str, torch.Tensor]) -> None:
        for l in range(self.layers):
            self.W[l].copy_(d[f"W{l}"])
            self.b[l].copy_(d[f"b{l}"])
        self.refresh_bf16()


def train(model: WarmStartMLP, ops: torch.Tensor, Y: torch.Tensor, label: torch.Tensor, epochs: int = 4,
          batch: int = 4096, lr: float = 1e-3, lr_final: float = 1e-5, seed: int = 0, log=None,
          val: Optional[Tuple[torch.Tensor, torch.Tensor, torch.Tensor]] = None) -> List[float]:
    """Adam over shuffled mini-batches with a cosine learning-rate decay from
    lr to lr_final.  Returns the mean training loss of every epoch; with
    val = (Ops, Y, label) the held-out loss of every epoch is appended to
    model.val_history."""
    dev = model.device
    P = ops.shape[0]
    model.val_history = []
    if P < batch:
        batch = max(32, P // 32 * 32)
    steps_per_epoch = max(1, P // batch)
    total = epochs * steps_per_epoch
    gen = torch.Generator(device=dev).manual_seed(int(seed))
    hist = []
    t = 0
    for ep in range(epochs):
        perm = torch.randperm(P, generator=gen, device=dev)
        acc = torch.zeros((), dtype=torch.float64, device=dev)
        for i in range(steps_per_epoch):
            idx = perm[i * batch:(i + 1) * batch]
            cur = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * t / max(1, total)))
            rows = model.train_step(ops[idx].contiguous(), Y[idx].contiguous(), label[idx].contiguous(), cur)
            acc += rows.double().sum()
            t += 1
        hist.append(float(acc) / (steps_per_epoch * batch))
        if val is not None:
            vl = sum(float(model.loss_rows(val[0][i:i + 16384].contiguous(), val[1][i:i + 16384].contiguous(),
                                           val[2][i:i + 16384].contiguous()).double().sum())
                     for i in range(0, val[0].shape[0], 16384)) / max(1, val[0].shape[0])
            model.val_history.append(vl)
        if log:
            log(f"epoch {ep + 1}/{epochs}: loss {hist[-1]:.4f}" +
                (f", held-out {model.val_history[-1]:.4f}" if val is not None else ""))
    return hist


def select_start(X, Y, Ops, Ch, C_warm, C_rand) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
    """Per problem, the start with the lower loss (warm start unless its loss
    is NaN or above the random start's).  Returns (C_start, used_warm (B,)
    int32, loss_warm, loss_rand); losses are FP64 means over the points."""
    B, M = Ops.shape
    # both starts in one evaluation call (2B expressions)
    both = fused_root(X, torch.cat([Ops, Ops]), torch.cat([Ch, Ch]), torch.cat([C_warm, C_rand]), torch.cat([Y, Y]),
                      mode=1, fwd_only=True, raise_errors=False)["loss"]
    lw, lr = both[:B], both[B:]
    C = torch.empty((B, M), dtype=torch.float32, device=Ops.device)
    used = torch.empty(B, dtype=torch.int32, device=Ops.device)
    _ck(_lib.hip().sr_mlp_select_start(_lib.ptr(C_warm), _lib.ptr(C_rand), _lib.ptr(lw), _lib.ptr(lr), M, B,
                                       _lib.ptr(C), _lib.ptr(used), _lib.stream_of(Ops.device)),
        "sr_mlp_select_start")
    return C, used, lw, lr


def warm_start(model: WarmStartMLP, X, Y, Ops, Ch, C_rand):
    """Network prediction with the random start as fallback (see select_start)."""
    C_warm = model.predict(Ops, Y)
    return select_start(X, Y, Ops, Ch, C_warm, C_rand)


def rel_mse_fp64(X, Y, Ops, Ch, C) -> torch.Tensor:
    """sum_n (f(X; C) - Y)^2 / sum_n (Y - mean Y)^2 per problem, f evaluated in
    FP64 at the FP32 constants (sum Y^2 as denominator when Y is constant;
    NaN or inf -> inf)."""
    y = fused_root(X.double(), Ops, Ch, C.double(), None, fwd_only=True, raise_errors=False)["y"]
    Yd = Y.double()
    num = ((y - Yd) ** 2).sum(dim=1)
    den = ((Yd - Yd.mean(dim=1, keepdim=True)) ** 2).sum(dim=1)
    den = torch.where(den > 0, den, (Yd ** 2).sum(dim=1))
    r = num / den
    return torch.where(torch.isfinite(r), r, torch.full_like(r, float("inf")))


def sample_problems(grammar, X: torch.Tensor, P: int, seed: int, c_std: float = 2.0, y_max: float = 1e6):
    """P fitting problems: expressions sampled from `grammar` with at least one
    constant, true constants ~ N(0, c_std^2), Y = f(X; C*) evaluated in FP64
    and rounded to FP32.  Problems whose Y is not finite, exceeds y_max in
    magnitude or is (nearly) constant are redrawn.  Returns device tensors
    (Ops, Ch, C_true, Y)."""
    dev = X.device
    gen = torch.Generator().manual_seed(int(seed))
    torch.manual_seed(int(seed))
    ops_l, ch_l, C_l, Y_l = [], [], [], []
    have = 0
    X64 = X.double()
    while have < P:
        ops, ch = grammar.sample(max(2 * (P - have), 1024))
        keep = (ops == int(Operator.LEARNABLE_CONSTANT)).any(dim=1)
        ops, ch = ops[keep], ch[keep]
        Ct = torch.zeros(ops.shape, dtype=torch.float32)
        m = ops == int(Operator.LEARNABLE_CONSTANT)
        Ct[m] = torch.randn(int(m.sum()), generator=gen) * c_std
        ops_d, ch_d, Ct_d = ops.to(dev), ch.to(dev), Ct.to(dev)
        y = fused_root(X64, ops_d, ch_d, Ct_d.double(), None, fwd_only=True, raise_errors=False)["y"].float()
        ok = torch.isfinite(y).all(dim=1) & (y.abs().amax(dim=1) < y_max) & (y.double().var(dim=1) > 1e-8)
        ops_l.append(ops_d[ok])
        ch_l.append(ch_d[ok])
        C_l.append(Ct_d[ok])
        Y_l.append(y[ok])
        have += int(ok.sum())
    cat = lambda l: torch.cat(l)[:P].contiguous()
    return cat(ops_l), cat(ch_l), cat(C_l), cat(Y_l)


def random_start(Ops: torch.Tensor, seed: int) -> torch.Tensor:
    """randn on the constant slots, 0 elsewhere (create_constants' 'randn')."""
    gen = torch.Generator(device=Ops.device).manual_seed(int(seed))
    C = torch.zeros(Ops.shape, dtype=torch.float32, device=Ops.device)
    m = Ops == int(Operator.LEARNABLE_CONSTANT)
    C[m] = torch.randn(int(m.sum()), generator=gen, device=Ops.device)
    return C


def make_training_set(grammar, X: torch.Tensor, P: int, seed: int, chunk: int = 65536, max_iter: int = 100,
                      label_max: float = 10.0, log=None):
    """(Ops, Y, label) tuples labelled by the plain L-BFGS: P problems are
    sampled, fitted from randn starts, and kept where the fit reached
    rel-MSE <= 1e-6 with every constant within +-label_max.  Returns
    (Ops, Ch, Y, label, stats)."""
    from .optimization import _fit

    t0 = time.time()
    ops, ch, Ct, Y = sample_problems(grammar, X, P, seed)
    C0 = random_start(ops, seed + 1)
    torch.cuda.synchronize(X.device)
    t1 = time.time()
    keep_l, C_l = [], []
    for i in range(0, P, chunk):
        sl = slice(i, min(P, i + chunk))
        C_fit, *_ = _fit(X, Y[sl], ops[sl], ch[sl], C0[sl], max_iter, 10, 1e-7, "strong_wolfe")
        r = rel_mse_fp64(X, Y[sl], ops[sl], ch[sl], C_fit)
        keep_l.append((r <= 1e-6) & (C_fit.abs().amax(dim=1) <= label_max))
        C_l.append(C_fit)
    keep = torch.cat(keep_l)
    C_fit = torch.cat(C_l)
    torch.cuda.synchronize(X.device)
    t2 = time.time()
    stats = dict(sampled=P, labelled=int(keep.sum()), sample_s=round(t1 - t0, 2), label_fit_s=round(t2 - t1, 2))
    if log:
        log(f"training set: {stats}")
    return ops[keep].contiguous(), ch[keep].contiguous(), Y[keep].contiguous(), C_fit[keep].contiguous(), stats

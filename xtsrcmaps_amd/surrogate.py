"""Training of the far-pixel cubature surrogate (the network of
csrc/xts_mlp.hpp) with the library's own HIP kernels: samples from the
device's FP64 cubature of an image (xts_surrogate_samples_dev), Adam on
minibatches (xts_surrogate_train_dev), then the packed BF16 network handed
to the context (xts_ctx_set_surrogate) for XTS_SURROGATE model maps.
torch only provides the device buffers; every computation is a library
kernel.  No counterpart in the reference."""
from __future__ import annotations

import dataclasses
import time

import numpy as np

from . import capi

H, IN = 64, 4


def init_params(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights, zero biases, in the layout of xts_mlp.hpp."""
    rng = np.random.default_rng(seed)
    p = np.zeros(capi.MLP_PARAMS, np.float32)
    off = 0
    for fo, fi in ((H, IN), (H, H), (H, H), (1, H)):
        a = np.sqrt(6.0 / (fo + fi))
        p[off:off + fo * fi] = rng.uniform(-a, a, fo * fi)
        off += fo * fi + (fo if fo > 1 else 1)
    return p


@dataclasses.dataclass
class Surrogate:
    params: object          # torch float32 [8705] on the device
    norm: object            # torch float32 [10] on the device
    n_samples: int
    steps: int
    final_loss: float
    seconds: dict


def samples(ctx, wcs, rule, n_samples: int, seed: int = 0, chunk: int = 1 << 24,
            max_chunks: int = 64):
    """n_samples training rows [n, 8] for the image and rule (rule [N, 3],
    the model map's), drawn in chunks of candidates until enough are found."""
    import torch
    dev = torch.device(f"cuda:{ctx.device}")
    r = torch.as_tensor(np.ascontiguousarray(rule, np.float64), device=dev)
    out = torch.empty((n_samples, 8), dtype=torch.float32, device=dev)
    count = torch.zeros(1, dtype=torch.int64, device=dev)
    # torch filled r and count on its own stream; the context may use another
    torch.cuda.synchronize(dev)
    for k in range(max_chunks):
        ctx.surrogate_samples_dev(wcs, r, seed * 1000003 + k, chunk, out, count)
        torch.cuda.current_stream().synchronize()
        ctx.synchronize()
        if int(count.item()) >= n_samples:
            return out
    raise capi.XtsError(1, f"only {int(count.item())} of {n_samples} surrogate samples in "
                           f"{max_chunks} chunks of {chunk} candidates")


def normalisation(rows: np.ndarray) -> np.ndarray:
    """mu [4], 1/sd [4], y scale, y offset from sample rows [n, 8]."""
    x = rows[:, :4].astype(np.float64)
    mu, sd = x.mean(0), np.maximum(x.std(0), 1e-6)
    ysc = max(float(rows[:, 4].std()), 1e-6)
    return np.concatenate([mu, 1.0 / sd, [ysc, 0.0]]).astype(np.float32)


def fit(ctx, wcs, rule, n_samples: int = 4_000_000, steps: int = 6000, batch: int = 8192,
        lr: float = 3e-3, seed: int = 0) -> Surrogate:
    """Train the network for the image and rule on the device and set it on
    the context."""
    import torch
    dev = torch.device(f"cuda:{ctx.device}")
    t0 = time.perf_counter()
    rows = samples(ctx, wcs, rule, n_samples, seed)
    t1 = time.perf_counter()
    host = rows[: min(n_samples, 1 << 20)].cpu().numpy()
    norm = torch.as_tensor(normalisation(host), device=dev)
    params = torch.as_tensor(init_params(seed), device=dev)
    hist = torch.zeros(steps, dtype=torch.float32, device=dev)
    torch.cuda.synchronize(dev)
    ctx.surrogate_train_dev(params, norm, rows, steps, batch, lr, seed, hist)
    ctx.synchronize()
    t2 = time.perf_counter()
    ctx.set_surrogate(params, norm)
    final = float(hist[-max(1, steps // 50):].mean().item())
    return Surrogate(params, norm, n_samples, steps, final,
                     {"samples": t1 - t0, "train": t2 - t1})

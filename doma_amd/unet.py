"""The learned source-iteration correction: network definition, parameter
layout, input features, and the trainer.

The product path evaluates the network with the hand-written BF16 MFMA kernels
of the native library (``doma_set_learned_correction``); this module is the
tooling around it.  ``UNet`` is the same architecture in PyTorch, the FP32
reference that the native kernels are tested against.  Parameters cross the
boundary as one flat FP32 vector, per layer W[cout][tap][cin] (tap = (dz+1)*9 +
(dy+1)*3 + (dx+1)) then bias[cout], in the layer order below.

Input channels, per cell, for group g: the last two increments of plain source
iterations d_k and d_(k-1), both divided by s = rms(d_k) over the grid, log
sigma_t, and log(1 / (1 - c + 1e-4)) with c the in-group scattering ratio
sigma_s,gg / sigma_t.  The remaining error is estimated as phi_converged - phi_k
= g (d_k + s y) with g = rho / (1 - rho), rho = rms(d_k) / rms(d_(k-1)) clamped
to 0.999: the geometric-series extrapolation g d_k, exact when one mode
dominates the iteration, plus the network's output y, which learns what the
extrapolation misses (spatially varying decay where slow modes of different
rates sit in different regions).  A zero network is the bare extrapolation.
"""
from __future__ import annotations

import math

import numpy as np

C1, C2, C3, CIN = 12, 24, 32, 4
RHO_MAX = 0.999
# (cin_a, cin_b, cout, taps): cin_b channels come from the coarser level, upsampled.
SPEC = [(CIN, 0, C1, 27), (C1, 0, C1, 27), (C1, 0, C2, 27), (C2, 0, C2, 27), (C2, 0, C3, 27),
        (C3, 0, C3, 27), (C2, C3, C2, 27), (C2, 0, C2, 27), (C1, C2, C1, 27), (C1, 0, C1, 27),
        (C1, 0, 1, 1)]


def param_count() -> int:
    return sum(co * t * (a + b) + co for a, b, co, t in SPEC)


def gain(dlast, dprev):
    """(s, g): s = rms(d_k), g = rho / (1 - rho) with rho = rms(d_k) / rms(d_(k-1))."""
    d = np.asarray(dlast, dtype=np.float64)
    dp = np.asarray(dprev, dtype=np.float64)
    s = math.sqrt(float((d * d).mean())) or 1.0
    sp = math.sqrt(float((dp * dp).mean()))
    rho = min(max(s / sp if sp > 0 else 0.0, 0.0), RHO_MAX)
    return s, rho / (1.0 - rho)


def features(dlast, dprev, sigt, cgg):
    """Network input [D, H, W, 4] from the last two increments and the group's
    cross sections; returns (features, s, g)."""
    s, g = gain(dlast, dprev)
    f = np.stack([np.asarray(dlast, dtype=np.float64) / s,
                  np.asarray(dprev, dtype=np.float64) / s,
                  np.log(np.maximum(sigt, 1e-6)), np.log(1.0 / (1.0 - cgg + 1e-4))], axis=-1)
    return f.astype(np.float32), s, g


def _torch():
    import torch
    return torch


def make_model():
    torch = _torch()
    nn = torch.nn
    F = torch.nn.functional

    class UNet(nn.Module):
        def __init__(self):
            super().__init__()
            self.convs = nn.ModuleList(
                [nn.Conv3d(a + b, co, 3 if t == 27 else 1, padding=1 if t == 27 else 0)
                 for a, b, co, t in SPEC])

        def forward(self, x):
            c = self.convs
            r = F.relu
            e1 = r(c[1](r(c[0](x))))
            e2 = r(c[3](r(c[2](F.avg_pool3d(e1, 2)))))
            b = r(c[5](r(c[4](F.avg_pool3d(e2, 2)))))
            d2 = torch.cat([e2, F.interpolate(b, scale_factor=2, mode="nearest")], dim=1)
            d2 = r(c[7](r(c[6](d2))))
            d1 = torch.cat([e1, F.interpolate(d2, scale_factor=2, mode="nearest")], dim=1)
            d1 = r(c[9](r(c[8](d1))))
            return c[10](d1)

    return UNet()


def flatten(model) -> np.ndarray:
    out = []
    for conv in model.convs:
        w = conv.weight.detach().cpu().float()          # [co][ci][kd][kh][kw]
        co, ci = w.shape[0], w.shape[1]
        out.append(w.permute(0, 2, 3, 4, 1).reshape(co, -1, ci).reshape(-1).numpy())
        out.append(conv.bias.detach().cpu().float().numpy())
    flat = np.concatenate(out).astype(np.float32)
    assert flat.size == param_count()
    return flat


def unflatten(model, flat):
    torch = _torch()
    off = 0
    with torch.no_grad():
        for conv, (a, b, co, t) in zip(model.convs, SPEC):
            ci = a + b
            k = 3 if t == 27 else 1
            n = co * t * ci
            w = torch.from_numpy(np.asarray(flat[off:off + n], dtype=np.float32))
            conv.weight.copy_(w.reshape(co, k, k, k, ci).permute(0, 4, 1, 2, 3))
            off += n
            conv.bias.copy_(torch.from_numpy(np.asarray(flat[off:off + co], dtype=np.float32)))
            off += co
    assert off == param_count()
    return model


def bf16_round(flat):
    """Round FP32 values to BF16 (nearest even), as the native kernels store them."""
    u = np.asarray(flat, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def collect_samples(batch, problems, max_per_group=None):
    """(features [D,H,W,4], scaled target [D,H,W]) pairs from a solved Batch: the
    target is what the extrapolation misses, (phi_converged - phi_k) / (s g) -
    d_k / s."""
    out = []
    for i, p in enumerate(problems):
        n = p.nx
        for g in range(p.groups):
            its = batch.sample_iterations(i, g)
            ks = range(len(its)) if max_per_group is None else range(min(len(its), max_per_group))
            st = p.sigt[p.mat, g]
            cgg = np.where(st > 0, p.sigs[p.mat, g, g] / np.where(st > 0, st, 1.0), 0.0)
            for k in ks:
                r, rp, t = batch.get_sample(i, g, k, with_prev=True)
                f, s, gn = features(r.reshape(n, n, n), rp.reshape(n, n, n), st.reshape(n, n, n),
                                    cgg.reshape(n, n, n))
                y = t.reshape(n, n, n) / (s * gn if gn > 0 else s) - f[..., 0]
                out.append((f, y.astype(np.float32)))
    return out


def train(samples, steps=2000, crops=(16, 32), batch_size=8, lr=1e-3, seed=0, device="cuda",
          log=None, holdout=0.05):
    """Adam on the relative L2 loss over random cubic crops (edge drawn from
    `crops` per step, samples at least that large); returns the flat parameters
    and the mean relative L2 error on held-out whole samples."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    torch.manual_seed(seed)
    order = rng.permutation(len(samples))
    nh = max(1, int(holdout * len(samples)))
    held = [samples[i] for i in order[:nh]]
    fit = [samples[i] for i in order[nh:]]
    model = make_model().to(device)
    opt = torch.optim.Adam(model.parameters(), lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, steps)
    xs = [torch.from_numpy(f).permute(3, 0, 1, 2).contiguous() for f, _ in fit]
    ys = [torch.from_numpy(t)[None].contiguous() for _, t in fit]
    sizes = np.array([x.shape[1] for x in xs])
    for step in range(steps):
        c = int(rng.choice(crops))
        cand = np.nonzero(sizes >= c)[0]
        if cand.size == 0:
            continue
        xb, yb = [], []
        for _ in range(batch_size):
            j = int(cand[rng.integers(cand.size)])
            n = xs[j].shape[1]
            o = [int(rng.integers(0, n - c + 1)) for _ in range(3)]
            xb.append(xs[j][:, o[0]:o[0] + c, o[1]:o[1] + c, o[2]:o[2] + c])
            yb.append(ys[j][:, o[0]:o[0] + c, o[1]:o[1] + c, o[2]:o[2] + c])
        x = torch.stack(xb).to(device)
        y = torch.stack(yb).to(device)
        pred = model(x)
        # relative L2 over the whole batch (no per-crop ratios of tiny norms)
        loss = ((pred - y) ** 2).sum() / ((y ** 2).sum() + 1e-12)
        opt.zero_grad()
        loss.backward()
        torch.nn.utils.clip_grad_norm_(model.parameters(), 1.0)
        opt.step()
        sched.step()
        if log and (step % 200 == 0 or step == steps - 1):
            log(f"train step {step} crop {c} loss {loss.item():.4f}")
    errs = []
    model.eval()
    with torch.no_grad():
        for f, t in held[:64]:
            x = torch.from_numpy(f).permute(3, 0, 1, 2)[None].to(device)
            pr = model(x)[0, 0].cpu().numpy()
            errs.append(float(np.linalg.norm(pr - t) / (np.linalg.norm(t) + 1e-30)))
    return flatten(model.cpu()), float(np.mean(errs)) if errs else float("nan")


def extrapolation_params() -> np.ndarray:
    """The ablation: a zero network, whose correction is the bare geometric-series
    extrapolation g d_k -- what isolates the trained network's contribution."""
    return np.zeros(param_count(), dtype=np.float32)


def scale_refinement(params, beta: float) -> np.ndarray:
    """Scale the network's output (the 1x1x1 head) by beta: the correction becomes
    g (d_k + beta s y).  beta is chosen on held-out solves; 0 is the bare
    extrapolation."""
    flat = np.array(params, dtype=np.float32, copy=True)
    flat[param_count() - (C1 + 1):] *= beta
    return flat


def init_params(seed=0, zero_head=False):
    """Fresh parameters without PyTorch: per layer W, b ~ U(-1/sqrt(fan_in),
    1/sqrt(fan_in)) (the default initialisation of a PyTorch Conv3d); with
    zero_head the 1x1x1 head starts at zero, so training starts from the bare
    extrapolation."""
    rng = np.random.default_rng(seed)
    out = []
    for a, b, co, t in SPEC:
        fan_in = (a + b) * t
        bound = 1.0 / math.sqrt(fan_in)
        w = rng.uniform(-bound, bound, co * t * (a + b))
        bias = rng.uniform(-bound, bound, co)
        if zero_head and t == 1:
            w[:] = 0.0
            bias[:] = 0.0
        out.append(w)
        out.append(bias)
    return np.concatenate(out).astype(np.float32)


def train_native(samples, steps=3000, crop=32, batch_size=8, lr=1e-3, seed=0, log=None,
                 holdout=0.05, params=None):
    """Training on the native hand-written kernels (BF16 MFMA forward and data
    gradients, FP32 weight gradients, Adam): random crops of edge `crop` from
    samples at least that large, cosine learning-rate decay.  Returns the flat
    parameters and the held-out relative L2 error (native forward pass)."""
    from . import capi

    rng = np.random.default_rng(seed)
    order = rng.permutation(len(samples))
    nh = max(1, int(holdout * len(samples)))
    held = [samples[i] for i in order[:nh]]
    fit = [samples[i] for i in order[nh:] if samples[i][0].shape[0] >= crop]
    if not fit:
        raise ValueError(f"no training sample is at least {crop} cells wide")
    flat = (init_params(seed, zero_head=True) if params is None
            else np.asarray(params, dtype=np.float32))
    c = crop
    with capi.NetTrainer(flat, crop=c, batch=batch_size) as tr:
        for step in range(steps):
            xb = np.empty((batch_size, c, c, c, CIN), dtype=np.float32)
            yb = np.empty((batch_size, c, c, c), dtype=np.float32)
            for k in range(batch_size):
                f, t = fit[int(rng.integers(len(fit)))]
                n = f.shape[0]
                o = [int(rng.integers(0, n - c + 1)) for _ in range(3)]
                xb[k] = f[o[0]:o[0] + c, o[1]:o[1] + c, o[2]:o[2] + c]
                yb[k] = t[o[0]:o[0] + c, o[1]:o[1] + c, o[2]:o[2] + c]
            cur = 0.5 * lr * (1.0 + math.cos(math.pi * step / steps))
            loss = tr.step(xb, yb.reshape(batch_size, -1), cur)
            if log and (step % 200 == 0 or step == steps - 1):
                log(f"native train step {step} loss {loss:.4f}")
        flat = tr.params()
    errs = []
    for f, t in held[:64]:
        if min(f.shape[:3]) % 4:
            continue
        pr = capi.net_forward(flat, f)
        errs.append(float(np.linalg.norm(pr - t) / (np.linalg.norm(t) + 1e-30)))
    return flat, float(np.mean(errs)) if errs else float("nan")


def training_problems(seed=20261210, small=False):
    """The batched data-generation set: uneven cubes (32, 48, 64 cells of 1 cm,
    8^3-cell blocks) of high-scattering media in the regime the correction
    serves -- sigma_t log-uniform in [0.5, 10], c ~ U[c_lo, 0.995] per (block,
    group) with c_lo in {0.9, 0.95}, 98 % of it in-group, 1 or 2 groups,
    sources in 5 or 10 % of the blocks -- diamond difference, 256 directions
    (8 polar x 4 azimuthal per octant, the learned workload's quadrature)."""
    from . import capi, workloads
    from .capi import MediaSpec

    rng = np.random.default_rng(seed)
    probs = []
    for b in range(96 // (4 if small else 1)):
        n = int(rng.choice([32, 48, 64]))
        groups = int(rng.choice([1, 2]))
        c_lo = float(rng.choice([0.9, 0.95]))
        frac = float(rng.choice([0.05, 0.1]))
        spec = MediaSpec(n=n, block=8, groups=groups, sigt_log=1, sigt_lo=0.5, sigt_hi=10.0,
                         c_lo=c_lo, c_hi=0.995, scatter_rule=capi.SCATTER_INGROUP,
                         source_rule=capi.SOURCE_BLOCKS, src_lo=0, src_hi=0,
                         source_fraction=frac, cell_width=1.0, seed=seed * 1000 + b)
        probs.append(workloads.block_media_problem(n, groups, 8, 4, capi.SCHEME_DD, spec,
                                                   f"train_{b}_{n}c_{groups}g"))
    return probs


def generate_samples(probs, fs_tol=1e-10, sample_every=5, sample_cap=16, log=None):
    """Run the problems as one batch (one persistent sweep launch per round) to a
    tight tolerance, so the targets phi_converged - phi_k are accurate, and
    collect the snapshots."""
    from .capi import Batch

    samples = []
    groups = {}
    for p in probs:  # one batch per quadrature (a batch shares one)
        groups.setdefault(p.dirs, []).append(p)
    for dirs, ps in groups.items():
        with Batch(ps) as b:
            b.autotune()
            stats, info = b.solve(fs_tol, 20000, fs_tol, 4, sample_every=sample_every,
                                  sample_cap=sample_cap)
            ok = sum(1 for s in stats if s["success"])
            if log:
                log(f"data: {len(ps)} problems, {info['rounds']} rounds, {ok} converged, "
                    f"{info['samples']} samples, {info['elapsed_ms'] / 1e3:.1f} s")
            samples += collect_samples(b, ps)
    return samples


def save_samples(path, samples):
    arrs = {}
    for i, (f, t) in enumerate(samples):
        arrs[f"f{i}"] = f
        arrs[f"t{i}"] = t
    np.savez(path, **arrs)


def load_samples(path):
    z = np.load(path, allow_pickle=False)
    n = len(z.files) // 2
    return [(z[f"f{i}"], z[f"t{i}"]) for i in range(n)]


def main(argv=None):
    """Training data and weights for the correction:
      python -m doma_amd.unet fit OUT.npy [--steps N] [--small]
      python -m doma_amd.unet data SAMPLES.npz [--small] [--sample-cap K]
      python -m doma_amd.unet train SAMPLES.npz OUT.npy [--steps N] [--torch]
    fit generates the data with the batched solver and trains with the library's
    own kernels in one process.  --torch trains the PyTorch twin instead; data and
    PyTorch training must then run in separate processes (the library and
    PyTorch each load their own HIP runtime, and only one of them can own the GPU
    in a process)."""
    import argparse
    import sys
    import time

    ap = argparse.ArgumentParser()
    ap.add_argument("phase", choices=["fit", "data", "train"])
    ap.add_argument("path")
    ap.add_argument("out", nargs="?")
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--small", action="store_true")
    ap.add_argument("--sample-cap", type=int, default=16)
    ap.add_argument("--torch", action="store_true",
                    help="train the PyTorch twin instead of the library's kernels")
    ap.add_argument("--beta", type=float, default=1.0,
                    help="scale of the trained refinement in the saved weights")
    args = ap.parse_args(argv)

    def log(m):
        print(m, file=sys.stderr, flush=True)

    t0 = time.time()
    if args.phase in ("fit", "data"):
        samples = generate_samples(training_problems(small=args.small),
                                   sample_cap=args.sample_cap, log=log)
        log(f"data: {len(samples)} pairs in {time.time() - t0:.0f} s")
        if args.phase == "data":
            save_samples(args.path, samples)
            log(f"data: written to {args.path}")
            return
        out = args.path
    else:
        samples = load_samples(args.path)
        out = args.out
        log(f"train: {len(samples)} pairs")
    if args.torch:
        flat, held_err = train(samples, steps=args.steps, log=log)
    else:
        flat, held_err = train_native(samples, steps=args.steps, log=log)
    np.save(out, scale_refinement(flat, args.beta))
    log(f"held-out relative L2 error of the predicted refinement: {held_err:.3f}"
        + (f"; saved with the refinement scaled by {args.beta}" if args.beta != 1.0 else ""))
    log(f"saved {out} ({flat.size} parameters) after {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()

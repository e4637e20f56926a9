"""Train the KPCNN denoiser on the path tracer's own renders.

Data: seeded scenes from the generators (sphere fields under the HDR sky seen
from random cameras, Cornell boxes with random sphere materials), each
rendered by the HIP library twice -- a noisy input at one of `--spp-in`
samples (colour, albedo, normal, depth AOVs) and a converged target at
`--spp-ref` samples.  Training: PyTorch on the GPU, float32, L1 loss on log(1 + colour)
over random crops, Adam.  Evaluation: held-out scenes, RMSE of log(1 + colour)
of the plain input, of the denoised input, and of plain renders at more
samples, to find the sample count plain accumulation needs to match the
denoiser ("equal RMSE").

    python -m optixrender_amd.train_denoiser --out optixrender_amd/denoiser_kpcnn.npz

Writes the weights (float32 masters, per layer w[co][ci][3][3] and b[co]) and
the evaluation as meta_* entries of the .npz, and prints one progress line
per step group so a long run shows it is alive.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import abi, denoise, scenes
from .renderer import Context, SceneData


def log(msg):
    print(f"[train] {msg}", flush=True)


def random_camera(rng, dist=(22.0, 34.0)):
    """A camera on a sphere around the sphere field, looking near its centre."""
    d = rng.normal(size=3)
    d[1] = abs(d[1]) * 0.6
    d /= np.linalg.norm(d)
    eye = d * rng.uniform(*dist)
    at = rng.uniform(-3.0, 3.0, 3)
    return abi.camera(tuple(eye), tuple(at))


def cornell_variant(rng) -> SceneData:
    sc = scenes.cornell_spheres(subdiv=3)
    kinds = [abi.CONDUCTOR, abi.DIELECTRIC, abi.PLASTIC, abi.DIFFUSE]
    for i in (3, 4):
        k = kinds[int(rng.integers(0, len(kinds)))]
        rough = float(rng.uniform(0.05, 0.8))
        col = tuple(rng.uniform(0.2, 0.9, 3))
        if k == abi.DIELECTRIC:
            sc.materials[i] = abi.material(k, roughness=rough, int_ior=1.5, ext_ior=1.0)
        elif k == abi.PLASTIC:
            sc.materials[i] = abi.material(k, roughness=rough, albedo=col, int_ior=1.5, ext_ior=1.0)
        elif k == abi.DIFFUSE:
            sc.materials[i] = abi.material(k, roughness=0.0, albedo=col)
        else:
            sc.materials[i] = abi.material(k, roughness=rough)
    return sc


def make_scene(seed: int, rng) -> SceneData:
    if seed % 4 == 3:
        return cornell_variant(rng)
    sc = scenes.sphere_field(seed=seed + 100, subdiv=4, env_w=1024, env_h=512)
    sc.camera = random_camera(rng)
    return sc


def render_aovs(ctx: Context, sid: int, sc: SceneData, w: int, h: int, spp: int, depth: int):
    ctx.select(sid)
    ctx.resize(w, h)
    ctx.set_camera(sc.camera)
    ctx.render(spp=spp, max_depth=depth, rr_start=0 if sc.env is not None else 3)
    return (ctx.download_aov(0, w, h), ctx.download_aov(1, w, h), ctx.download_aov(2, w, h),
            ctx.download_depth(w, h))


def dataset(ctx, seeds, w, h, spp_in, spp_ref, depth):
    """Per scene: features (10, h, w), noisy colour (3, h, w), target (3, h, w).
    spp_in: the input sample counts, used in turn over the scenes."""
    out = []
    for i, seed in enumerate(seeds):
        spp = spp_in[i % len(spp_in)]
        rng = np.random.default_rng(seed)
        sc = make_scene(seed, rng)
        sid = ctx.upload(sc)
        try:
            ctx.build_bvh(sid)
            c, a, n, d = render_aovs(ctx, sid, sc, w, h, spp, depth)
            ref, _, _, _ = render_aovs(ctx, sid, sc, w, h, spp_ref, depth)
        finally:
            ctx.release(sid)
        feat = denoise.features(c, a, n, d).transpose(2, 0, 1)
        out.append((feat, c[..., :3].transpose(2, 0, 1).copy(), ref[..., :3].transpose(2, 0, 1).copy(),
                    (c, a, n, d, ref)))
        if (i + 1) % 8 == 0 or i + 1 == len(seeds):
            log(f"rendered {i + 1}/{len(seeds)} scenes")
    return out


def lrmse(a, b):
    return denoise.log_rmse(a, b)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(denoise.WEIGHTS))
    ap.add_argument("--scenes", type=int, default=128)
    ap.add_argument("--eval-scenes", type=int, default=6)
    ap.add_argument("--res", type=int, default=256)
    ap.add_argument("--spp-in", default="2,4,8", help="input sample counts, used in turn over the training scenes")
    ap.add_argument("--eval-spp", type=int, default=4)
    ap.add_argument("--spp-ref", type=int, default=4096)
    ap.add_argument("--depth", type=int, default=16)
    ap.add_argument("--iters", type=int, default=12000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--crop", type=int, default=128)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)

    import torch

    torch.manual_seed(args.seed)
    # PyTorch's own im2col convolution: no MIOpen kernel compilation on a fresh machine
    torch.backends.cudnn.enabled = False
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    ctx = Context(0)
    train_seeds = list(range(args.scenes))
    eval_seeds = list(range(1000, 1000 + args.eval_scenes))
    spp_in = [int(x) for x in args.spp_in.split(",")]
    train = dataset(ctx, train_seeds, args.res, args.res, spp_in, args.spp_ref, args.depth)
    held = dataset(ctx, eval_seeds, args.res, args.res, [args.eval_spp], args.spp_ref, args.depth)
    log(f"data ready in {time.perf_counter() - t0:.1f}s")

    feats = torch.from_numpy(np.stack([f for f, _, _, _ in train])).to(dev)
    noisy = torch.from_numpy(np.stack([c for _, c, _, _ in train])).to(dev)
    target = torch.from_numpy(np.stack([r for _, _, r, _ in train])).to(dev)
    model = denoise.torch_model().to(dev)
    opt = torch.optim.Adam(model.parameters(), lr=args.lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, T_max=args.iters)
    n, _, H, W = feats.shape
    cr = args.crop
    g = torch.Generator(device="cpu").manual_seed(args.seed)
    t1 = time.perf_counter()
    for it in range(args.iters):
        idx = torch.randint(0, n, (args.batch,), generator=g)
        ys = torch.randint(0, H - cr + 1, (args.batch,), generator=g)
        xs = torch.randint(0, W - cr + 1, (args.batch,), generator=g)
        fb = torch.stack([feats[i, :, y:y + cr, x:x + cr] for i, y, x in zip(idx, ys, xs)])
        cb = torch.stack([noisy[i, :, y:y + cr, x:x + cr] for i, y, x in zip(idx, ys, xs)])
        tb = torch.stack([target[i, :, y:y + cr, x:x + cr] for i, y, x in zip(idx, ys, xs)])
        # the device kernel runs on BF16 weights and activations: train with
        # float32 masters, the rounding is a small perturbation
        out = model(fb, cb)
        loss = torch.nn.functional.l1_loss(torch.log1p(out.clamp_min(0)), torch.log1p(tb))
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        sched.step()
        if (it + 1) % 1000 == 0 or it == 0:
            log(f"step {it + 1}/{args.iters} loss {loss.item():.5f} ({time.perf_counter() - t1:.1f}s)")

    layers = denoise.torch_layers(model)
    # held-out evaluation with the BF16 reference of the device kernel
    ev = {"plain": [], "denoised": []}
    for _, _, _, (c, a, nrm, d, ref) in held:
        dn = denoise.reference_denoise(layers, c, a, nrm, d, bf16=True)
        ev["plain"].append(lrmse(c, ref))
        ev["denoised"].append(lrmse(dn, ref))
    meta = {
        "train_scenes": args.scenes, "eval_scenes": args.eval_scenes, "res": args.res, "spp_in": spp_in,
        "eval_spp": args.eval_spp,
        "spp_ref": args.spp_ref, "iters": args.iters, "batch": args.batch, "crop": args.crop,
        "eval_log_rmse_plain": float(np.mean(ev["plain"])), "eval_log_rmse_denoised": float(np.mean(ev["denoised"])),
        "seconds": round(time.perf_counter() - t0, 1),
    }
    log("eval " + json.dumps(meta))
    denoise.save_weights(layers, args.out, meta={"json": json.dumps(meta)})
    log(f"wrote {args.out}")
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Train the KPCNN denoiser on the path tracer's own renders.

Data: seeded scenes of the headline generator (`--generator sphere_field`:
12 icospheres at subdivision 6 with seeded centres, radii and materials under
the 2048x1024 HDR sky, seen from random cameras; `mixed` adds Cornell boxes
with random sphere materials), each rendered by the HIP library twice -- a
noisy input at one of `--spp-in` samples (colour, albedo, normal, depth AOVs)
and a converged target at `--spp-ref` samples.  Training: PyTorch on the
GPU, float32, L1 loss on log(1 + colour) over random crops, Adam.

Evaluation: `--eval-scenes` held-out seeds at `--eval-res`, the input at
`--eval-spp` samples denoised by the library's own BF16 kernel, against a
`--spp-ref` plain render: RMSE of the displayed image (the tone map in float)
and of log(1 + colour), beside plain accumulation at the input's and at
more samples, and the plain sample count that matches the denoiser's
display RMSE (log-log interpolation).

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


def make_scene(seed: int, rng, generator: str = "sphere_field") -> SceneData:
    if generator == "mixed" and seed % 4 == 3:
        return cornell_variant(rng)
    sc = scenes.sphere_field(seed=seed + 100, subdiv=6)
    sc.camera = random_camera(rng)
    return sc


def render_aovs(ctx: Context, sid: int, sc: SceneData, w: int, h: int, spp: int, depth: int):
    ctx.select(sid)
    ctx.resize(w, h)
    ctx.set_camera(sc.camera)
    ctx.render(spp=spp, max_depth=depth, rr_start=0 if sc.env is not None else 3)
    return (ctx.download_aov(0, w, h), ctx.download_aov(1, w, h), ctx.download_aov(2, w, h),
            ctx.download_depth(w, h))


def dataset(ctx, seeds, w, h, spp_in, spp_ref, depth, generator):
    """Per scene: features (10, h, w), noisy colour (3, h, w), target (3, h, w).
    spp_in: the input sample counts, used in turn over the scenes."""
    out = []
    for i, seed in enumerate(seeds):
        spp = spp_in[i % len(spp_in)]
        rng = np.random.default_rng(seed)
        sc = make_scene(seed, rng, generator)
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


def evaluate(ctx, layers, seeds, w, h, spp, spp_ref, depth, generator, plain_spp=(4, 16, 64, 256)):
    """Held-out scenes at w x h: the library's denoiser on spp-sample inputs
    against spp_ref plain renders, beside plain accumulation."""
    from .serve_bench import display_rmse

    ctx.load_denoiser(denoise.flat_weights(layers))
    rows = []
    for seed in seeds:
        rng = np.random.default_rng(seed)
        sc = make_scene(seed, rng, generator)
        sid = ctx.upload(sc)
        try:
            ctx.build_bvh(sid)
            ref = render_aovs(ctx, sid, sc, w, h, spp_ref, depth)[0]
            c, a, n, d = render_aovs(ctx, sid, sc, w, h, spp, depth)
            dn = ctx.denoise(c, a, n, d)
            plain = {s: display_rmse(render_aovs(ctx, sid, sc, w, h, s, depth)[0], ref) for s in plain_spp}
        finally:
            ctx.release(sid)
        rows.append({"seed": seed, "display_rmse_denoised": display_rmse(dn, ref),
                     "display_rmse_plain": plain, "log_rmse_denoised": lrmse(dn, ref),
                     "log_rmse_plain": lrmse(c, ref)})
        log(f"eval seed {seed}: display RMSE denoised {rows[-1]['display_rmse_denoised']:.4f}, plain "
            f"{ {k: round(v, 4) for k, v in plain.items()} }")
    dn = float(np.mean([r["display_rmse_denoised"] for r in rows]))
    pl = {s: float(np.mean([r["display_rmse_plain"][s] for r in rows])) for s in plain_spp}
    xs, ys = np.log(np.array(plain_spp, float)), np.log(np.array([pl[s] for s in plain_spp]))
    if math.log(dn) <= ys[-1]:
        equal = f">{plain_spp[-1]}"
    elif math.log(dn) >= ys[0]:
        equal = f"<={plain_spp[0]}"
    else:
        i = int(np.searchsorted(-ys, -math.log(dn)))
        f = (math.log(dn) - ys[i - 1]) / (ys[i] - ys[i - 1])
        equal = round(float(math.exp(xs[i - 1] + f * (xs[i] - xs[i - 1]))), 1)
    return {"eval_res": [w, h], "eval_seeds": list(seeds), "eval_spp": spp,
            "eval_display_rmse_denoised": dn, "eval_display_rmse_plain": {str(k): v for k, v in pl.items()},
            "eval_plain_spp_at_equal_display_rmse": equal,
            "eval_log_rmse_plain": float(np.mean([r["log_rmse_plain"] for r in rows])),
            "eval_log_rmse_denoised": float(np.mean([r["log_rmse_denoised"] for r in rows]))}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(denoise.WEIGHTS))
    ap.add_argument("--generator", choices=["sphere_field", "mixed"], default="sphere_field")
    ap.add_argument("--scenes", type=int, default=200)
    ap.add_argument("--eval-scenes", type=int, default=20)
    ap.add_argument("--res", type=int, default=512)
    ap.add_argument("--eval-res", default="1920x1080")
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
    train = dataset(ctx, train_seeds, args.res, args.res, spp_in, args.spp_ref, args.depth, args.generator)
    log(f"data ready in {time.perf_counter() - t0:.1f}s")

    feats = torch.from_numpy(np.stack([f for f, _, _, _ in train])).to(dev)
    noisy = torch.from_numpy(np.stack([c for _, c, _, _ in train])).to(dev)
    target = torch.from_numpy(np.stack([r for _, _, r, _ in train])).to(dev)
    del train
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
    del feats, noisy, target

    layers = denoise.torch_layers(model)
    ew, eh = (int(v) for v in args.eval_res.split("x"))
    ev = evaluate(ctx, layers, eval_seeds, ew, eh, args.eval_spp, args.spp_ref, args.depth, args.generator)
    meta = {
        "generator": args.generator, "train_scenes": args.scenes, "res": args.res, "spp_in": spp_in,
        "spp_ref": args.spp_ref, "iters": args.iters, "batch": args.batch, "crop": args.crop, **ev,
        "seconds": round(time.perf_counter() - t0, 1),
    }
    log("eval " + json.dumps(meta))
    denoise.save_weights(layers, args.out, meta={"json": json.dumps(meta)})
    log(f"wrote {args.out}")
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

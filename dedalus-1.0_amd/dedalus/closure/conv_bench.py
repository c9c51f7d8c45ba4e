"""Time the closure network's convolutions alone (for kernel profiles).

    python -m dedalus.closure.conv_bench [--batch 256] [--reps 20]

Runs the 5-layer width-32 network (3 -> 32 -> 32 -> 32 -> 32 -> 2 channels)
on a batch of 64^2 images: the preallocated inference pass and one training
forward + backward, and prints one JSON line with the mean times and rates.
"""

import argparse
import json

import torch

from .net import ClosureNet, Inference


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--size", type=int, default=64)
    ap.add_argument("--reps", type=int, default=20)
    args = ap.parse_args()
    B, H = args.batch, args.size
    net = ClosureNet(cin=3, cout=2, width=32, depth=5, seed=0).cuda()
    inf = Inference(net, B, H, H)
    inf.x.normal_()
    x = torch.randn(B, H, H, 3, device="cuda")
    inf.run()
    net(x).square().mean().backward()
    torch.cuda.synchronize()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    e[0].record()
    for _ in range(args.reps):
        inf.run()
    e[1].record()
    for _ in range(args.reps):
        net.zero_grad(set_to_none=True)
        net(x).square().mean().backward()
    e[2].record()
    e[2].synchronize()
    fwd_ms = e[0].elapsed_time(e[1]) / args.reps
    train_ms = e[1].elapsed_time(e[2]) / args.reps
    ch = net.chans
    alg = sum(2 * 9 * ch[l] * ch[l + 1] for l in range(len(ch) - 1)) * B * H * H
    print(json.dumps({"batch": B, "size": H, "forward_ms": round(fwd_ms, 4),
                      "forward_TFLOPs": round(alg / fwd_ms / 1e9, 2),
                      "train_step_ms": round(train_ms, 4),
                      "train_TFLOPs": round(3 * alg / train_ms / 1e9, 2)}), flush=True)


if __name__ == "__main__":
    main()

"""Autotuner for the traversal launch shape of a tracer or scene batch.

Searches persistent blocks per CU (waves per CU = 4 x blocks), the LDS stack
ring (16 / 24 / 32 entries per lane) and the traversal steps between per-lane
ray refills (how often lanes whose rays finished are compacted away and
refilled), timing whole frames with HIP events.  Results never depend on the
launch shape, only the time does.
"""

from __future__ import annotations

import itertools
import json
import time

import torch

DEFAULT_SPACE = {
    "blocks_per_cu": (0, 4),
    "lds_stack": (16, 24, 32),
    "refill_steps": (8, 16, 32),
}


def autotune(target, command_encoder, *, space=None, frames: int = 2, start_iteration: int = 1_000_000,
             log=None) -> dict:
    """Times every candidate on ``target`` (a TextureSpacePathTracer with a
    payload bound via ``run``, or a SceneBatch) and leaves the fastest set.

    ``target.execute(encoder, iteration, ...)`` must be callable with only
    the encoder and an iteration (SceneBatch), so tracers are wrapped by
    :func:`tracer_target`.  Iterations far from the caller's are used; the
    caller resets its caches afterwards."""
    space = space or DEFAULT_SPACE
    stream = torch.cuda.current_stream()
    results = []
    it = start_iteration
    for bpc, ring, steps in itertools.product(space["blocks_per_cu"], space["lds_stack"], space["refill_steps"]):
        target.tune(bpc, ring, steps)
        target.execute(command_encoder, it)  # allocation + warm caches
        it += 1
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record(stream)
        for _ in range(frames):
            target.execute(command_encoder, it)
            it += 1
        end.record(stream)
        torch.cuda.synchronize()
        ms = start.elapsed_time(end) / frames
        results.append({"blocks_per_cu": bpc, "lds_stack": ring, "refill_steps": steps, "ms_per_frame": ms})
        if log:
            log(f"autotune {bpc=} {ring=} {steps=}: {ms:.3f} ms/frame")
    best = min(results, key=lambda r: r["ms_per_frame"])
    target.tune(best["blocks_per_cu"], best["lds_stack"], best["refill_steps"])
    return {"best": best, "results": results}


class tracer_target:
    """Adapts a TextureSpacePathTracer and its payload to the autotuner."""

    def __init__(self, tracer, payload):
        self.tracer = tracer
        self.payload = payload

    def tune(self, *args):
        self.tracer.tune(*args)

    def execute(self, command_encoder, iteration):
        self.tracer.execute(command_encoder, self.payload, iteration)


def save(result: dict, path) -> None:
    with open(path, "w") as f:
        json.dump({"time": time.time(), **result}, f, indent=1)

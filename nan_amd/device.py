"""Device and command recording on top of PyTorch-ROCm.

The reference records passes into a SlangPy command encoder and submits
them once per frame (app.py:509-537).  Here a :class:`Device` owns one HIP
stream (torch's current stream unless another default-priority stream is
given) and a
:class:`CommandEncoder` launches each pass on that stream as it is
recorded; ``submit_command_buffer`` is the point where the host may wait.
Recording inside ``torch.cuda.graph(...)`` captures the launches into a
hipGraph, because every pass launches on the device's current stream.
"""

from __future__ import annotations

import torch


class Device:
    """One GPU plus the stream passes launch on."""

    def __init__(self, index: int = 0, stream: "torch.cuda.Stream | None" = None):
        if not torch.cuda.is_available():
            raise RuntimeError("nan_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.torch_device = torch.device("cuda", index)
        # torch's current stream by default, so tensors written by torch and
        # the passes launched here are ordered without extra events
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.torch_device)

    @property
    def stream_handle(self) -> int:
        """hipStream_t of the stream passes launch on (the current stream
        while a graph is being captured)."""
        current = torch.cuda.current_stream(self.torch_device)
        if torch.cuda.is_current_stream_capturing():
            return current.cuda_stream
        return self.stream.cuda_stream

    def create_command_encoder(self) -> "CommandEncoder":
        return CommandEncoder(self)

    def submit_command_buffer(self, command_buffer=None) -> None:
        """Work is already enqueued on the stream; nothing to flush."""

    def wait_for_idle(self) -> None:
        self.stream.synchronize()

    def empty(self, shape, dtype=torch.float32) -> torch.Tensor:
        with torch.cuda.stream(self.stream):
            return torch.empty(shape, dtype=dtype, device=self.torch_device)

    def zeros(self, shape, dtype=torch.float32) -> torch.Tensor:
        with torch.cuda.stream(self.stream):
            return torch.zeros(shape, dtype=dtype, device=self.torch_device)


class CommandEncoder:
    """Launches passes on the device stream as they are recorded."""

    def __init__(self, device: Device):
        self.device = device

    @property
    def stream_handle(self) -> int:
        return self.device.stream_handle

    def finish(self) -> "CommandEncoder":
        return self

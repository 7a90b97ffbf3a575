"""Side streams that run concurrently with the main stream."""
from __future__ import annotations

import time
from typing import Dict

# Two side streams per device and main stream for the whole process, for solves that overlap the
# main stream's work (the concurrent VUMPS site step).  HIP multiplexes streams over a few hardware queues, and work of streams that share a
# queue runs in order: a side stream on the main stream's queue (or both on one queue) serialises
# the concurrent solves (measured: D = 1024 VUMPS iterations 0.15 s instead of 0.11 s).  Which
# queue a stream of torch's pool lands on is not exposed, so the pair is chosen by measurement: a
# bounded spin kernel on the main stream and on each candidate together must take about as long
# as one alone.
_SIDE_STREAMS: Dict[tuple, list] = {}


def _span(torch, streams, cycles) -> float:
    """Wall time of one spin kernel of `cycles` clock cycles on each stream, issued together."""
    torch.cuda.synchronize()
    t = time.perf_counter()
    for st in streams:
        with torch.cuda.stream(st):
            torch.cuda._sleep(cycles)
    torch.cuda.synchronize()
    return time.perf_counter() - t


def side_streams(torch, device: int) -> list:
    """The two side streams for torch's current stream on `device` (chosen on the first call)."""
    main = torch.cuda.current_stream(device)
    key = (device, main.cuda_stream)
    if key not in _SIDE_STREAMS and not hasattr(torch.cuda, "_sleep"):
        with torch.cuda.device(device):  # no spin kernel to measure with: two pool streams
            _SIDE_STREAMS[key] = [torch.cuda.Stream() for _ in range(2)]
    if key not in _SIDE_STREAMS:
        with torch.cuda.device(device):
            cycles = 1 << 18
            one = _span(torch, [main], cycles)  # also loads the kernel
            while one < 2e-3 and cycles < 1 << 30:
                cycles *= 4
                one = _span(torch, [main], cycles)
            picked, drawn = [], []
            for _ in range(16):
                st = torch.cuda.Stream()  # default priority
                drawn.append(st)
                if _span(torch, [main] + picked + [st], cycles) < 1.5 * one:
                    picked.append(st)
                    if len(picked) == 2:
                        break
            _SIDE_STREAMS[key] = (picked + [st for st in drawn if st not in picked])[:2]
    return _SIDE_STREAMS[key]

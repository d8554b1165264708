"""Serving workload: the request mix of the serving configuration and a load test of the
request server (``fft3d.Server``) against the plain one-launch-per-request path.

Mix (one draw per request): L uniform over {6, 8, 12, 17, 24, 36} plus the anisotropic
8x8x16; batch log-uniform in [1, 512]; direction forward/backward 50/50; complex128 with
probability 0.7, else complex64; inputs iid uniform(-1, 1), resident on the device before
the trace starts (one pool per geometry and precision, each request reading a random
window); arrivals Poisson at a fixed rate; admission capped by the server's budget.
"""
import math

import numpy as np

from . import fft3d, hip

GEOMS = [(6, 6, 6), (8, 8, 8), (12, 12, 12), (17, 17, 17), (24, 24, 24), (36, 36, 36),
         (8, 8, 16)]
MAX_BATCH = 512


def draw_mix(n, seed=0):
    """n requests: (shape, batch, direction, precision) drawn as described above."""
    rng = np.random.default_rng(seed)
    g = rng.integers(0, len(GEOMS), n)
    b = np.clip(np.floor(np.exp(rng.uniform(0.0, math.log(MAX_BATCH + 1), n))), 1, MAX_BATCH)
    d = np.where(rng.random(n) < 0.5, -1, 1)
    p = np.where(rng.random(n) < 0.7, fft3d.C128, fft3d.C64)
    return [(GEOMS[g[i]], int(b[i]), int(d[i]), int(p[i])) for i in range(n)]


class InputPools:
    """Device-resident uniform(-1,1) inputs: per geometry and precision, 2*MAX_BATCH
    volumes, so a request of batch b can read any window of b consecutive volumes.
    keep_host=True keeps the host copies (host(shape, prec)) for checking results."""

    def __init__(self, seed=0, keep_host=False):
        rng = np.random.default_rng(seed)
        self.bufs = {}
        self.hosts = {}
        for shape in GEOMS:
            V = shape[0] * shape[1] * shape[2]
            for prec in (fft3d.C128, fft3d.C64):
                dt = np.complex128 if prec == fft3d.C128 else np.complex64
                x = (rng.uniform(-1, 1, (2 * MAX_BATCH, V)) +
                     1j * rng.uniform(-1, 1, (2 * MAX_BATCH, V))).astype(dt)
                buf = hip.DeviceBuffer(x.nbytes)
                buf.upload(x)
                self.bufs[(shape, prec)] = (buf, x.itemsize * V)
                if keep_host:
                    self.hosts[(shape, prec)] = x

    def host(self, shape, prec):
        """Host copy of a pool, (2*MAX_BATCH, V) (needs keep_host=True)."""
        return self.hosts[(tuple(shape), prec)]

    def addr(self, shape, prec, first):
        buf, vb = self.bufs[(shape, prec)]
        return buf.addr + first * vb

    def close(self):
        for buf, _ in self.bufs.values():
            buf.close()


def make_requests(mix, pools, seed=1):
    rng = np.random.default_rng(seed)
    out = []
    for shape, b, d, p in mix:
        first = int(rng.integers(0, MAX_BATCH + 1))
        out.append(fft3d.SrvRequest(*shape, b, d, p, pools.addr(shape, p, first), None))
    return out


def poisson_arrivals(n, rate, seed=2):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, n))


def load_point(server, requests, rate, seed=2):
    """Replay at `rate` requests/s; p50/p99 latency (ms) and achieved throughput."""
    arr = poisson_arrivals(len(requests), rate, seed)
    lat = server.run_trace(requests, arr)
    span = float(arr[-1] + 0.0)
    xf = sum(r.batch for r in requests)
    return dict(rate=rate, p50_ms=float(np.percentile(lat, 50) * 1e3),
                p99_ms=float(np.percentile(lat, 99) * 1e3), mean_ms=float(lat.mean() * 1e3),
                max_ms=float(lat.max() * 1e3),
                req_per_s=len(requests) / max(span, 1e-9) if span > 0 else 0.0,
                transforms_per_s=xf / max(span, 1e-9))


def max_rate_under(server, requests, rates, p99_ms=2.0):
    """Walk `rates` upward; the last rate whose p99 stays within p99_ms, and every point."""
    points, best = [], None
    for r in rates:
        pt = load_point(server, requests, r)
        points.append(pt)
        if pt["p99_ms"] <= p99_ms:
            best = pt
        else:
            break
    return best, points

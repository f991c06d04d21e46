"""Synthetic workloads with known spectra, generated on the device.

* ``ReflectedGenerator`` writes A = (I - 2uu') [diag(sigma); 0] (I - 2vv') (m x n, n <= m) with
  random unit vectors u, v (``mcnla_synth_reflected``): dense, singular values exactly sigma, O(mn)
  work, so a serving client can make a fresh matrix per request.
* ``serving_stream`` draws the mixed-size request stream: m log-uniform in [1000, 50000], n
  log-uniform in [500, 5000] (resampled until n <= m), k uniform in {5..50}, p = 10, N uniform in
  {8..32}, sigma_i = rho^i with rho ~ U(0.8, 0.98), per-request matrix and Omega seeds.
* ``sign_block_stack`` builds a tall matrix whose row blocks are sign-flipped copies of one dense
  square block, scaled so that the block stack has orthonormal columns: its exact singular values
  are those of the square block, computable on the host.

No reference counterpart: the reference demo draws Haar factors on the host
(demo/demo.cpp:250-293).
"""
import ctypes

import numpy as np

from . import _lib
from .device import DeviceArray


class ReflectedGenerator:
    """Writes reflected matrices into caller-provided device buffers on its own stream.  The
    vectors (up to ``max_m + 2 max_n`` doubles per matrix) go through a ring of ``slots`` device
    staging buffers; a slot is reused only after the kernel that read it has finished (an event
    per slot), so ``write_async`` never waits unless the ring has wrapped onto unfinished work."""

    def __init__(self, max_m, max_n, slots=128):
        self.max_m, self.max_n = int(max_m), int(max_n)
        self.width = self.max_m + 2 * self.max_n
        self.slots = int(slots)
        self._stage = DeviceArray((self.slots, self.width), np.float64)
        L = _lib.lib()
        hp = ctypes.c_void_p()
        _lib.check(L.mcnla_host_malloc(ctypes.byref(hp), self.slots * self.width * 8))
        self._pinned_ptr = hp.value
        self._pinned = np.ctypeslib.as_array(
            ctypes.cast(hp.value, ctypes.POINTER(ctypes.c_double)),
            shape=(self.slots, self.width))
        h = ctypes.c_void_p()
        _lib.check(L.mcnla_stream_create(ctypes.byref(h)))
        self.stream = h.value
        self._events = []
        for _ in range(self.slots):
            e = ctypes.c_void_p()
            _lib.check(L.mcnla_event_create(ctypes.byref(e)))
            self._events.append(e.value)
        self._used = [False] * self.slots
        self._next = 0

    @staticmethod
    def vectors(m, n, seed):
        """The unit vectors u (m) and v (n) of matrix seed `seed`."""
        rng = np.random.default_rng(seed)
        u = rng.standard_normal(m)
        v = rng.standard_normal(n)
        return u / np.linalg.norm(u), v / np.linalg.norm(v)

    def write_async(self, out, sigma, seed, vectors=None):
        """Enqueues A into ``out`` (a DeviceArray, m x n with n <= m, any dtype/order/ld) on the
        generator's stream; returns an event that completes when A is written.  ``vectors``:
        the (u, v) of ``vectors(m, n, seed)``, if the caller computed them already."""
        m, n = out.shape
        assert n <= m <= self.max_m and n <= self.max_n and len(sigma) == n
        L = _lib.lib()
        slot = self._next
        self._next = (slot + 1) % self.slots
        ev = self._events[slot]
        if self._used[slot]:
            _lib.check(L.mcnla_event_synchronize(ev))
        u, v = self.vectors(m, n, seed) if vectors is None else vectors
        host = self._pinned[slot]  # free: the slot's last copy and kernel have finished
        host[:n] = sigma
        host[n:n + m] = u
        host[n + m:2 * n + m] = v
        base = self._stage.ptr + slot * self.width * 8
        _lib.check(L.mcnla_memcpy_htod_async(base, host.ctypes.data, (2 * n + m) * 8,
                                             self.stream))
        _lib.check(L.mcnla_synth_reflected(m, n, base, base + 8 * n, base + 8 * (n + m), out.ptr,
                                           out.ld, out.dtype_code, out.order_code, self.stream))
        _lib.check(L.mcnla_event_record(ev, self.stream))
        self._used[slot] = True
        return ev

    def write(self, out, sigma, seed):
        """As write_async, and waits for it."""
        _lib.check(_lib.lib().mcnla_event_synchronize(self.write_async(out, sigma, seed)))
        return out

    def make(self, m, n, sigma, seed, dtype=np.float32, order="F"):
        return self.write(DeviceArray((m, n), dtype, order), sigma, seed)

    @staticmethod
    def host(m, n, sigma, seed):
        """The same matrix in FP64 on the host (for checking on small sizes)."""
        u, v = ReflectedGenerator.vectors(m, n, seed)
        D = np.zeros((m, n))
        D[np.arange(n), np.arange(n)] = sigma
        Hu = np.eye(m) - 2.0 * np.outer(u, u)
        Hv = np.eye(n) - 2.0 * np.outer(v, v)
        return Hu @ D @ Hv

    def free(self):
        L = _lib.lib()
        if self.stream:
            _lib.check(L.mcnla_stream_synchronize(self.stream))
            for e in self._events:
                L.mcnla_event_destroy(e)
            L.mcnla_stream_destroy(self.stream)
            self.stream = None
            self._events = []
            self._pinned = None
            L.mcnla_host_free(self._pinned_ptr)
        self._stage.free()


def serving_stream(count, seed=3):
    """The request stream: a list of dicts (m, n, k, p, N, rho, a_seed, omega_seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        m = int(round(np.exp(rng.uniform(np.log(1000), np.log(50000)))))
        while True:
            n = int(round(np.exp(rng.uniform(np.log(500), np.log(5000)))))
            if n <= m:
                break
        k = int(rng.integers(5, 51))
        N = int(rng.integers(8, 33))
        rho = float(rng.uniform(0.8, 0.98))
        out.append({"m": m, "n": n, "k": k, "p": 10, "N": N, "rho": rho,
                    "a_seed": 100000 + i, "omega_seed": 5000 + i})
    return out


def stream_sigma(req):
    return req["rho"] ** np.arange(req["n"], dtype=np.float64)


def sign_block_stack(blocks, M, seed):
    """Host array A = [diag(s_0) M; diag(s_1) M; ...] (row-major, M's dtype; M is b x n), with
    s_t random sign vectors.  A = U (sqrt(blocks) M) where U = [diag(s_t)] / sqrt(blocks) has
    orthonormal columns, so the singular values of A are exactly sqrt(blocks) times those of M
    (the sign flips are exact in any dtype).  Returns A and the signs (blocks x b)."""
    rng = np.random.default_rng(seed)
    b, n = M.shape
    signs = rng.choice(np.array([-1.0, 1.0], dtype=M.dtype), size=(blocks, b))
    out = np.empty((blocks * b, n), dtype=M.dtype)
    for i in range(blocks):
        np.multiply(M, signs[i][:, None], out=out[i * b:(i + 1) * b])
    return out, signs

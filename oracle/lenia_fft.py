"""FP64 Lenia by FFT in numpy: the oracle for large periodic lattices.

TEST INFRASTRUCTURE ONLY.  Restates the FFTW extension's kernel field
(ext/CellularAutomataCoreFFTWExt.jl:17-30): the kernel is wrapped onto the
lattice at mod(o, dims) for every read offset o (kpad[mod1(o + 1, dims)] += w),
and the field is the cross-correlation irfft(rfft(src) .* conj(rfft(kpad))).
The growth is local_rule(::Lenia) for Lenia() (src/lenia.jl:42-47): FP64 on the
Float32 cell, clamped to [0, 1], rounded to Float32.
"""
import numpy as np


class LeniaFFT:
    def __init__(self, dims, taps, weights, mu, sigma, dt):
        kpad = np.zeros(dims, dtype=np.float64)
        for (o1, o2), w in zip(taps, weights):
            kpad[o1 % dims[0], o2 % dims[1]] += w
        self.kf = np.conj(np.fft.rfft2(kpad))
        self.dims = tuple(dims)
        self.mu, self.sigma, self.dt = float(mu), float(sigma), float(dt)

    def field(self, u):
        src = np.asarray(u, dtype=np.float64)
        return np.fft.irfft2(np.fft.rfft2(src) * self.kf, s=self.dims)

    def growth(self, u, U):
        d = U - self.mu
        g = 2.0 * np.exp(-(d * d) / (2.0 * (self.sigma * self.sigma))) - 1.0
        v = np.asarray(u, dtype=np.float64) + self.dt * g
        return np.clip(v, 0.0, 1.0).astype(np.float32)

    def step(self, u):
        return self.growth(u, self.field(u))

"""Seeded synthetic inputs for the search workloads.

Definitions (written here because the workloads only name them):
  * noise: spectra get complex Gaussian bins with E|X|^2 = 1 (re, im ~ N(0, 1/2)),
    the normalisation the analytic noise scale assumes; time series get N(0, 1)
    samples and are normalised by realfft + rednoise.
  * pulse profile: von Mises, p(phi) = exp(kappa (cos 2 pi phi - 1)), with kappa
    from the FWHM duty cycle w: kappa = ln 2 / (1 - cos(pi w)); its harmonic
    amplitudes are I_h(kappa) / I_0(kappa).
  * S/N: the optimal coherent matched-filter S/N over all injected harmonics,
    sqrt(2 sum_h |c_h|^2) for spectral amplitudes c_h in units where the noise has
    E|X|^2 = 1; for a time series sqrt(N) * rms(signal) / sigma_noise (the same
    number after rednoise).
  * a spectral line at fractional bin r contributes c * e^{i pi d} sin(pi d)/(pi d)
    to bin k, d = r - k (the DFT of a finite-length complex sinusoid), injected
    over +-64 bins.
"""
import numpy as np
from scipy import special

DT = 64e-6


def vonmises_harmonics(duty, nharm):
    kappa = np.log(2.0) / (1.0 - np.cos(np.pi * duty))
    h = np.arange(1, nharm + 1)
    return special.ive(h, kappa) / special.ive(0, kappa), kappa


def spectrum_with_pulsar(nbins, r0, nharm, duty, snr, seed, noise=True):
    """Complex64 spectrum of nbins bins, pulsar harmonics at h*r0."""
    rng = np.random.default_rng(seed)
    if noise:
        amps = ((rng.standard_normal(nbins) + 1j * rng.standard_normal(nbins))
                * np.sqrt(0.5)).astype(np.complex128)
    else:
        amps = np.zeros(nbins, dtype=np.complex128)
    if snr > 0 and r0 > 0:
        a, _ = vonmises_harmonics(duty, nharm)
        phases = rng.uniform(0, 2 * np.pi) * np.arange(1, nharm + 1)   # pulse at phi0
        scale = snr / np.sqrt(2.0 * np.sum(a ** 2))
        for h in range(1, nharm + 1):
            r = h * r0
            k0 = int(np.floor(r))
            ks = np.arange(max(k0 - 63, 1), min(k0 + 65, nbins))
            if ks.size == 0:
                continue
            d = r - ks
            resp = np.exp(1j * np.pi * d) * np.sinc(d)
            amps[ks] += scale * a[h - 1] * np.exp(1j * phases[h - 1]) * resp
    return amps.astype(np.complex64)


def config1(seed=1):
    """.fft of 2^20 bins (N = 2^21, dt 64 us), pulsar at r0 = 12345.37 with 60
    harmonics, duty 0.05, S/N 15."""
    nbins = 1 << 20
    amps = spectrum_with_pulsar(nbins, 12345.37, 60, 0.05, 15.0, seed)
    return amps, 2 * nbins, DT


def series_with_pulsar(N, period, duty, snr, seed, dt=DT):
    """float32 series: N(0,1) noise plus a von Mises pulsar of the given
    period (s); S/N = sqrt(N) * rms(signal)."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(N, dtype=np.float32)
    if snr > 0 and period > 0:
        kappa = np.log(2.0) / (1.0 - np.cos(np.pi * duty))
        phase0 = rng.uniform()
        t = np.arange(N, dtype=np.float64) * dt
        ph = np.mod(t / period + phase0, 1.0)
        s = np.exp(kappa * (np.cos(2 * np.pi * ph) - 1.0))
        s -= s.mean()
        s *= snr / (np.sqrt(N) * np.sqrt(np.mean(s ** 2)))
        x += s.astype(np.float32)
    return x


def config2(seed=2, N=1 << 24):
    """2^24 samples at 64 us with a 5.3 ms pulsar (duty 0.03, S/N 12)."""
    return series_with_pulsar(N, 5.3e-3, 0.03, 12.0, seed), DT


def smooth7(n):
    """Smallest even 7-smooth integer >= n."""
    m = int(n)
    while True:
        if m % 2 == 0:
            k = m
            for p in (2, 3, 5, 7):
                while k % p == 0:
                    k //= p
            if k == 1:
                return m
        m += 1


def dm_sweep(nfiles=64, seed=4, lo=22, hi=25, npulsar=8, same_length=None):
    """Lengths log-uniform in [2^lo, 2^hi] rounded up to even 7-smooth sizes
    (or all `same_length`, the DM trials of one observation); a 3.7 ms pulsar
    in the npulsar files around the middle DM with S/N
    18 * exp(-0.5 (d/2)^2) at DM offset d; nharms cycles 16, 32, 60."""
    rng = np.random.default_rng(seed)
    lens = [smooth7(int(2 ** rng.uniform(lo, hi))) for _ in range(nfiles)]
    if same_length:
        lens = [int(same_length)] * nfiles
    mid = nfiles // 2
    out = []
    for i, N in enumerate(lens):
        d = i - mid
        has = -(npulsar // 2) <= d < npulsar - npulsar // 2
        snr = 18.0 * np.exp(-0.5 * (d / 2.0) ** 2) if has else 0.0
        out.append(dict(N=N, seed=seed * 1000 + i, snr=snr, period=3.7e-3, duty=0.04,
                        nharms=(16, 32, 60)[i % 3]))
    return out


def serving_requests(n=500, rate=20.0, seed=5, lo=20, hi=25):
    """Poisson arrivals at `rate` per second; N log-uniform in [2^lo, 2^hi]
    (even 7-smooth); nharms uniform over 16, 32, 60; 10% carry a pulsar with
    period log-uniform in 1-500 ms and S/N uniform in 8-20."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    reqs = []
    for i in range(n):
        N = smooth7(int(2 ** rng.uniform(lo, hi)))
        puls = rng.uniform() < 0.1
        reqs.append(dict(arrival=float(t[i]), N=N, seed=seed * 100000 + i,
                         nharms=int(rng.choice([16, 32, 60])),
                         period=float(np.exp(rng.uniform(np.log(1e-3), np.log(0.5)))) if puls else 0.0,
                         snr=float(rng.uniform(8, 20)) if puls else 0.0, duty=0.05))
    return reqs


def filterbank_with_pulsar(nspec, nchan, lofreq, chan_wid, dt, dm, period, duty, snr, seed,
                           dtype=np.float32):
    """[nspec, nchan] spectra, lowest channel first: per-channel Gaussian noise
    plus a von Mises pulsar delayed in each channel by the cold-plasma law,
    DM / (0.000241 f^2) s with f in MHz, relative to the highest channel.  snr
    is the S/N of the series de-dispersed at `dm` (sqrt(nspec) * rms of the
    summed pulse over the summed noise).  float32 noise is N(0, 1); uint8 is
    round(96 + 8 N(0, 1) + pulse) clipped to [0, 255]."""
    rng = np.random.default_rng(seed)
    scale = 8.0 if dtype == np.uint8 else 1.0
    x = rng.standard_normal((nchan, nspec), dtype=np.float32)
    if dtype == np.uint8:
        x *= scale
        x += 96.0
    if snr > 0 and period > 0:
        kappa = np.log(2.0) / (1.0 - np.cos(np.pi * duty))
        nb = 4096
        prof = np.exp(kappa * (np.cos(2 * np.pi * (np.arange(nb) + 0.5) / nb) - 1.0))
        prof -= prof.mean()
        prof *= 1.0 / np.sqrt(np.mean(prof ** 2))
        amp = scale * snr / (np.sqrt(nspec) * np.sqrt(nchan))
        phase0 = rng.uniform()
        ftop = lofreq + (nchan - 1) * chan_wid
        t = np.arange(nspec, dtype=np.float64) * dt
        for c in range(nchan):
            f = lofreq + c * chan_wid
            delay = dm / (0.000241 * f * f) - dm / (0.000241 * ftop * ftop)
            ph = np.mod((t - delay) / period + phase0, 1.0)
            x[c] += (amp * prof[(ph * nb).astype(np.int64) % nb]).astype(np.float32)
    if dtype == np.uint8:
        x = np.clip(np.rint(x), 0, 255).astype(np.uint8)
    return np.ascontiguousarray(x.T)

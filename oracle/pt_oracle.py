"""CPU restatement of the device parallel-tempering step (TEST INFRASTRUCTURE: only tests/ import it).

The sampler rules follow src/mcmc_sampler_internals.cpp:
  * mcmc_step (:35-145) with the Gaussian step only, gaussian_step (:364-409): one dimension beta chosen
    uniformly, proposed[beta] = width[chain][beta] * N(0, 1) + current[beta];
  * the prior: a uniform box in the sampling basis; outside it the step is rejected without a likelihood
    (:83-87, lp = limit_inf); a NaN likelihood is rejected too (:93-95);
  * Metropolis-Hastings: reject iff (-ll + ll') / T < log U (:100, :118-119);
  * chain_swap (:1086-1115): for every ladder, pairs (k, k+1) in increasing k; single_chain_swap
    (:1121-1180): equal temperatures never swap, reject iff exp((ll1 - ll2) / T2 - (ll1 - ll2) / T1) < U.
Random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
SC'11), pinned by the Random123 known-answer vectors in tests/test_pt_oracle.py; the stream layout (which
counter feeds which draw) is the one documented in the engine's header for gwh_pt_create.
The likelihood is a callback, so the same restatement runs with the C oracle or any other evaluator.
"""
import numpy as np

_M32 = np.uint64(0xFFFFFFFF)
_MUL = (np.uint64(0xD2511F53), np.uint64(0xCD9E8D57))
_WEYL = (np.uint64(0x9E3779B9), np.uint64(0xBB67AE85))


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """vectorised Philox4x32-10 on arrays of 32-bit values held in uint64; returns four uint64 arrays"""
    c = [np.asarray(x, dtype=np.uint64) & _M32 for x in (c0, c1, c2, c3)]
    k = [np.uint64(int(k0) & 0xFFFFFFFF), np.uint64(int(k1) & 0xFFFFFFFF)]
    for _ in range(10):
        p0 = _MUL[0] * c[0]
        p1 = _MUL[1] * c[2]
        c = [(p1 >> np.uint64(32)) ^ c[1] ^ k[0], p1 & _M32, (p0 >> np.uint64(32)) ^ c[3] ^ k[1], p0 & _M32]
        k = [(k[0] + _WEYL[0]) & _M32, (k[1] + _WEYL[1]) & _M32]
    return c


def u53(a, b):
    """53-bit uniform in [0, 1) from two 32-bit words"""
    return ((a >> np.uint64(5)).astype(np.float64) * 67108864.0 + (b >> np.uint64(6)).astype(np.float64)) \
        / 9007199254740992.0


def mh_draws(seed, step, n, dim):
    """(dimension, normal deviate, log U) of every instance for one step"""
    i = np.arange(n, dtype=np.uint64)
    slo, shi = step & 0xFFFFFFFF, step >> 32
    k0, k1 = seed & 0xFFFFFFFF, seed >> 32
    A = philox4x32_10(i, slo, shi, 0, k0, k1)
    B = philox4x32_10(i, slo, shi, 1, k0, k1)
    beta = ((A[0] * np.uint64(dim)) >> np.uint64(32)).astype(np.int64)
    u1, u2 = u53(A[1], A[2]), u53(A[3], B[0])
    z = np.sqrt(-2.0 * np.log(1.0 - u1)) * np.cos(np.pi * (2.0 * u2))
    with np.errstate(divide="ignore"):
        logu = np.log(u53(B[1], B[2]))
    return beta, z, logu


def swap_draws(seed, step, nt, nc):
    """U of pair k of ladder c, shape (nt - 1, nc)"""
    if nt < 2:
        return np.zeros((0, nc))
    idx = np.arange((nt - 1) * nc, dtype=np.uint64)
    C = philox4x32_10(idx, step & 0xFFFFFFFF, step >> 32, 2, seed & 0xFFFFFFFF, seed >> 32)
    return u53(C[0], C[1]).reshape(nt - 1, nc)


class PTOracle:
    """n_temps x n_chains chains, instance i = temp * n_chains + chain; loglike(thetas) -> log-likelihoods"""

    def __init__(self, loglike, theta0, widths, temps, n_chains, lower, upper, seed, swap_every):
        self.loglike = loglike
        self.temps = np.asarray(temps, dtype=np.float64)
        self.nt, self.nc = len(self.temps), int(n_chains)
        self.n = self.nt * self.nc
        self.lower, self.upper = np.asarray(lower, float), np.asarray(upper, float)
        self.dim = len(self.lower)
        self.theta = np.array(np.broadcast_to(np.asarray(theta0, float), (self.n, self.dim)))
        self.width = np.array(np.broadcast_to(np.asarray(widths, float), (self.n, self.dim)))
        self.T = np.repeat(self.temps, self.nc)
        self.seed, self.swap_every = int(seed), int(swap_every)
        self.ll = np.asarray(self.loglike(self.theta), dtype=np.float64)
        self.step_no = 0
        self.history = []

    def step(self):
        s = self.step_no
        beta, z, logu = mh_draws(self.seed, s, self.n, self.dim)
        rows = np.arange(self.n)
        prop = self.theta.copy()
        prop[rows, beta] = self.width[rows, beta] * z + self.theta[rows, beta]
        ok = np.all((prop >= self.lower) & (prop <= self.upper), axis=1)
        llp = np.full(self.n, np.nan)
        if ok.any():
            llp[ok] = self.loglike(prop[ok])
        with np.errstate(invalid="ignore"):
            ratio = np.where(ok & ~np.isnan(llp), (-self.ll + llp) / self.T, -np.inf)
            accept = (ratio > -np.inf) & ~(ratio < logu)
        self.theta[accept] = prop[accept]
        self.ll[accept] = llp[accept]
        swapped = np.zeros((max(self.nt - 1, 0), self.nc), dtype=bool)
        if self.swap_every > 0 and (s + 1) % self.swap_every == 0:
            alpha = swap_draws(self.seed, s, self.nt, self.nc)
            for c in range(self.nc):
                for k in range(self.nt - 1):
                    T1, T2 = self.temps[k], self.temps[k + 1]
                    if T1 == T2:
                        continue
                    i1, i2 = k * self.nc + c, (k + 1) * self.nc + c
                    d = self.ll[i1] - self.ll[i2]
                    if not (np.exp(d / T2 - d / T1) < alpha[k, c]):
                        self.theta[[i1, i2]] = self.theta[[i2, i1]]
                        self.ll[[i1, i2]] = self.ll[[i2, i1]]
                        swapped[k, c] = True
        self.history.append({"prop": prop, "llp": llp, "accept": accept, "swapped": swapped})
        self.step_no += 1

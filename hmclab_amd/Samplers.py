"""GPU HMC sampler with the reference's ``HMC(seed).sample(...)`` interface.

Mirrors ``hmclab.Samplers.HMC`` (hmclab/Samplers.py:1105-1755) and the
proposal loop of ``_AbstractSampler`` (Samplers.py:321-704): argument checks
and messages, the samples-file layout and attributes, thinning, autotuning,
``max_time`` and ``KeyboardInterrupt`` handling.  The trajectory, the
Hamiltonians and the Metropolis test run on the GPU for a whole batch of
``chains`` at once (the reference runs one chain per process,
Samplers.py:1816-1977).

Random numbers:
* ``rng="host"`` (default for one chain) draws, per chain and per proposal,
  ``normal(size=(d,1))`` (scaled by sqrt(M) for a diagonal mass), then
  ``uniform(0.5,1.5)`` when ``randomize_stepsize``, then ``uniform(0,1)`` from
  ``numpy.random.default_rng(seed)`` — the reference's draw order
  (MassMatrices.py:142,227; Samplers.py:1533-1534,1486) — so chain c
  reproduces a reference ``HMC(seed=seeds[c])`` chain up to floating-point
  rounding.  Seeds default to ``seed + c``.
* ``rng="device"`` (default for several chains) draws the same quantities on
  the GPU with a Philox4x32-10 counter RNG keyed by ``seed``; whole chunks of
  proposals then run without host round trips.

Learned surrogate (``surrogate=``, a trained :class:`hmclab_amd.surrogate.Surrogate`
of the likelihood misfit): its gradient drives the trajectory while every
Metropolis test uses the exact FP64 Hamiltonian.  When the acceptance rate of
a chunk of proposals (device RNG) or of the last ``surrogate_window``
proposals (host RNG) falls below ``surrogate_fallback``, the sampler detaches
the surrogate and continues with the exact gradient;
``surrogate_fallback_proposal`` records where (None if it never happened).
"""
from __future__ import annotations

import math
import time
import warnings
from datetime import datetime

import numpy as np

from . import _lib
from .Distributions import _AbstractDistribution, build_problem
from .engine import Chains
from .MassMatrices import Unit, _AbstractMassMatrix
from .Samples import Samples, SamplesWriter


class HMC:
    """Hamiltonian Monte Carlo on the GPU (Samplers.py:1105-1178)."""

    name = "Hamiltonian Monte Carlo"
    stepsize: float = 0.1
    amount_of_steps: int = 10
    minimal_stepsize: float = 1e-18
    integrators_full_names = {
        "lf": "leapfrog integrator",
        "3s": "three stage integrator",
        "4s": "four stage integrator",
    }
    available_integrators = integrators_full_names.keys()

    def __init__(self, seed: int = None):
        self.seed = seed
        self.rng = np.random.default_rng() if seed is None else np.random.default_rng(seed)
        self.times_started = 0
        self.diagnostics = None

    # ------------------------------------------------------------------ #
    def sample(self, samples_filename, distribution, stepsize: float = 0.1,
               randomize_stepsize: bool = True, amount_of_steps: int = 10, mass_matrix=None,
               integrator: str = "lf", initial_model=None, proposals: int = 100,
               online_thinning: int = 1, diagnostic_mode: bool = False,
               overwrite_existing_file: bool = False, max_time: float = None,
               autotuning: bool = False, target_acceptance_rate: float = 0.65,
               learning_rate: float = 0.75, queue=None, disable_progressbar=False, *,
               chains: int = 1, precision: str = "f64", engine: str = "auto", rng: str = None,
               seeds=None, chunk: int = None, surrogate=None, surrogate_fallback: float = 0.5,
               surrogate_window: int = 20, temperatures=None, exchange_schedule=None,
               exchange_interval: int = 1, rngs=None):
        """Sample ``distribution``; see the module docstring for the extras.

        ``samples_filename``: a ``.npy`` path (one chain), a list of paths (one
        per chain), a path used as a pattern ``<stem>_<c>.npy`` for several
        chains, or None to keep the samples in memory (``self.results``,
        shape (chains, d+1, N)).
        """
        self._temperatures = temperatures
        self._given_rngs = rngs
        self._init(samples_filename, distribution, stepsize, randomize_stepsize, amount_of_steps,
                   mass_matrix, integrator, initial_model, proposals, online_thinning,
                   diagnostic_mode, overwrite_existing_file, max_time, autotuning,
                   target_acceptance_rate, learning_rate, disable_progressbar, chains, precision,
                   engine, rng, seeds, chunk)
        try:
            self._setup_exchange(exchange_schedule, exchange_interval)
            self._attach_surrogate(surrogate, surrogate_fallback, surrogate_window)
            self._loop()
        finally:
            self._close()
        return self

    # ------------------------------------------------------------------ #
    def _init(self, samples_filename, distribution, stepsize, randomize_stepsize, amount_of_steps,
              mass_matrix, integrator, initial_model, proposals, online_thinning, diagnostic_mode,
              overwrite_existing_file, max_time, autotuning, target_acceptance_rate, learning_rate,
              disable_progressbar, chains, precision, engine, rng, seeds, chunk):
        # Samplers.py:347-416
        assert isinstance(distribution, _AbstractDistribution), (
            "The passed target distribution should be a derived class of _AbstractDistribution."
        )
        assert type(distribution.dimensions) == int and distribution.dimensions > 0, (
            "The passed target distribution should have an integer dimension larger than zero."
        )
        self.distribution = distribution
        self.dimensions = d = distribution.dimensions
        assert type(proposals) == int and proposals > 0, (
            "The amount of proposal requested (`proposals`) should be an integer number larger than zero."
        )
        assert type(online_thinning) == int and online_thinning > 0, (
            "The amount of online thinning (`online_thinning`) should be an integer number larger than zero."
        )
        assert proposals % online_thinning == 0, (
            "The amount of proposals (`proposals`) needs to be a multiple of the online thinning "
            "(`online_thinning`) number, to prevent sample wastage."
        )
        assert type(chains) == int and chains > 0, "`chains` should be a positive integer."
        self.proposals = proposals
        self.online_thinning = online_thinning
        self.proposals_after_thinning = proposals // online_thinning
        self.chains = B = chains
        # Samplers.py:1334-1384
        if autotuning:
            assert 0.5 < learning_rate <= 1.0, (
                "The learning rate should be larger than 0.5 and smaller than or equal to 1.0, "
                f"otherwise the Markov chain does not converge. Chosen: {learning_rate}"
            )
        self.stepsize = float(stepsize)
        assert self.stepsize > 0.0, "Stepsize should be a float larger than zero."
        assert type(amount_of_steps) == int, (
            "The amount of steps (amount_of_steps) the HMC integrator should make should be an integer."
        )
        assert amount_of_steps > 0, (
            "The amount of steps (amount_of_steps) the HMC integrator should make should be larger than zero."
        )
        self.amount_of_steps = amount_of_steps
        self.randomize_stepsize = bool(randomize_stepsize)
        self.autotuning = bool(autotuning)
        self.target_acceptance_rate = float(target_acceptance_rate)
        self.learning_rate = float(learning_rate)
        if mass_matrix is None:
            mass_matrix = Unit(d)
        assert isinstance(mass_matrix, _AbstractMassMatrix), (
            "The passed mass matrix (mass_matrix) should be a class derived from _AbstractMassMatrix."
        )
        assert mass_matrix.dimensions == d, (
            "The passed mass matrix (mass_matrix) should have dimensions equal to the target "
            f"distribution. Passed: {mass_matrix.dimensions}, required: {d}."
        )
        mass_matrix.rng = self.rng
        self.mass_matrix = mass_matrix
        self.integrator = str(integrator)
        if self.integrator not in self.available_integrators:
            raise ValueError(f"Unknown integrator used. Choices are: {self.available_integrators}")
        if max_time is not None:
            max_time = float(max_time)
            assert max_time > 0.0, "The maximal runtime (`max_time`) should be a float larger than zero."
        self.max_time = max_time
        self.diagnostic_mode = bool(diagnostic_mode)
        self.disable_progressbar = disable_progressbar
        self.rng_mode = rng or ("host" if B == 1 else "device")
        assert self.rng_mode in ("host", "device"), "rng must be 'host' or 'device'"

        # initial models (Samplers.py:406-426)
        if initial_model is None:
            X0 = np.zeros((B, d))
        else:
            im = np.array(initial_model, dtype=np.float64)
            if im.size == d:
                X0 = np.broadcast_to(im.reshape(1, d), (B, d)).copy()
            else:
                assert im.shape == (d, B), (
                    "The initial model (`initial_model`) dimension is incompatible with the target "
                    f"distribution. Supplied model shape: {im.shape}.Required shape: {(d, 1)}"
                )
                X0 = np.ascontiguousarray(im.T)

        # files
        self.samples_filename = samples_filename
        self._writers = []
        if samples_filename is not None:
            if isinstance(samples_filename, (list, tuple)):
                names = list(samples_filename)
                assert len(names) == B, "one samples filename per chain is required"
            else:
                assert type(samples_filename) == str, (
                    "First argument should be a string containing the path of the file to which to "
                    f"write samples. It was an object of type {type(samples_filename)}."
                )
                if B == 1:
                    names = [samples_filename]
                else:
                    stem, ext = samples_filename.rsplit(".", 1) if "." in samples_filename else (samples_filename, "npy")
                    names = [f"{stem}_{c}.{ext}" for c in range(B)]
            self._filenames = names
            for nme in names:
                self._writers.append(SamplesWriter(nme, d, overwrite=bool(overwrite_existing_file)))
        self._memory = [] if samples_filename is None else None

        # device objects
        self.problem = build_problem(distribution, precision=precision, engine=engine,
                                     mass_diagonal=mass_matrix._device_diagonal())
        self._chains = Chains(self.problem, B)
        T = self._temperatures if self._temperatures is not None else distribution.temperature
        T = np.broadcast_to(np.asarray(T, dtype=np.float64), (B,))
        if np.any(T != 1.0):
            self._chains.set_temperatures(T)
        self.temperatures = T.copy()
        try:
            self._chains.set_state(X0)
        except ValueError as e:
            self._close_files()
            raise AssertionError(str(e)) from None
        X, chi = self._chains.get_state()
        self.current_model = X[0].reshape(d, 1) if B == 1 else X.T.copy()
        self.current_x = float(chi[0]) if B == 1 else chi
        self.accepted_proposals = 0 if B == 1 else np.zeros(B, dtype=np.int64)
        self.amount_of_writes = 0
        self.current_proposal = 0

        # per-chain randoms for host mode
        if self.rng_mode == "host":
            if self._given_rngs is not None:
                assert len(self._given_rngs) == B, "one random generator per chain is required"
                self._rngs = list(self._given_rngs)
            elif B == 1 and seeds is None:
                self._rngs = [self.rng]
            else:
                if seeds is None:
                    base = 0 if self.seed is None else int(self.seed)
                    seeds = [base + c for c in range(B)]
                assert len(seeds) == B, "one seed per chain is required"
                self._rngs = [np.random.default_rng(s) for s in seeds]
        self._device_seed = int(self.seed) if self.seed is not None else int(self.rng.integers(0, 2**63))
        self._step = np.full(B, self.stepsize)
        if self.autotuning:
            self.acceptance_rates = np.empty((proposals, B))
            self.stepsizes = np.empty((proposals, B))
            if self.rng_mode == "device":
                # the device records exp(H0 - H1) and the pre-update step of every
                # proposal, as HMC.autotune does (Samplers.py:1496-1497)
                self._chains.set_tuning_history(proposals)
        self._chunk = chunk
        if self.diagnostic_mode:
            self._chains.enable_timing(True)

        self.start_time = datetime.now()
        self.times_started += 1
        attrs = dict(write_index=-1, last_written_sample=-1, proposals=proposals, acceptance_rate=0,
                     online_thinning=online_thinning,
                     start_time=self.start_time.strftime("%d-%b-%Y (%H:%M:%S.%f)"), sampler=self.name,
                     stepsize=self.stepsize, amount_of_steps=amount_of_steps,
                     mass_matrix=mass_matrix.name, integrator=self.integrators_full_names[self.integrator])
        for w in self._writers:
            for k, v in attrs.items():
                if k not in ("write_index", "last_written_sample"):
                    w.write_attribute(k, v)

    # ------------------------------------------------------------------ #
    def _setup_exchange(self, schedule, interval):
        """Replica exchange between chains (ParallelSampleSMP, Samplers.py:589-669):
        after proposal i with i % interval == 0, row i // interval of the schedule
        pairs its entries (2k, 2k+1); the second of a pair draws the uniform."""
        self._ex_schedule = None
        if schedule is None:
            return
        sched = np.asarray(schedule, dtype=np.int32)
        assert sched.ndim == 2 and sched.shape[1] % 2 == 0, "exchange schedule must be (events, 2 * pairs)"
        assert int(interval) > 0, "exchange_interval must be positive"
        self._ex_schedule = sched
        self._ex_interval = int(interval)
        if self.rng_mode == "device":
            self._chains.set_exchange(self._ex_interval, sched)

    def _host_exchange(self, i):
        if self._ex_schedule is None or i % self._ex_interval != 0:
            return
        e = i // self._ex_interval
        if e >= self._ex_schedule.shape[0]:
            return
        pairs = self._ex_schedule[e].reshape(-1, 2)
        u = np.array([self._rngs[int(b)].uniform(0, 1) for _, b in pairs])
        self._chains.exchange(pairs, u, i)

    def _attach_surrogate(self, surrogate, fallback, window):
        self.surrogate = surrogate
        self.surrogate_fallback = float(fallback)
        self.surrogate_fallback_proposal = None
        self._sur_window = int(window)
        self._sur_recent = []
        if surrogate is not None:
            assert 0.0 <= self.surrogate_fallback < 1.0, "surrogate_fallback must be in [0, 1)"
            assert self._sur_window > 0, "surrogate_window must be positive"
            self._chains.set_surrogate(surrogate)

    def _surrogate_check(self, accepted_fraction, proposals_done):
        """Fallback to the exact gradient when the acceptance rate is too low."""
        if self.surrogate is None or self.surrogate_fallback_proposal is not None:
            return
        if accepted_fraction < self.surrogate_fallback:
            self._chains.set_surrogate(None)
            self.surrogate_fallback_proposal = int(proposals_done)
            warnings.warn(f"surrogate acceptance {accepted_fraction:.3f} < {self.surrogate_fallback}: "
                          f"continuing with the exact gradient after proposal {proposals_done}")

    # ------------------------------------------------------------------ #
    def _write(self, X, chi):
        """Record one thinned sample per chain: X (B, d), chi (B,)."""
        cols = np.concatenate([X, chi[:, None]], axis=1)  # (B, d+1)
        if self._writers:
            for c, w in enumerate(self._writers):
                w.append(cols[c][:, None])
        else:
            self._memory.append(cols.copy())
        self.amount_of_writes += 1

    def _progress(self):
        try:
            import tqdm.auto as tq
            return tq.tqdm(total=self.proposals, desc="Tot. acc rate: 0.00. Progress", leave=True,
                           disable=self.disable_progressbar)
        except Exception:   # pragma: no cover
            return None

    def _loop(self):
        B = self.chains
        deadline = None if self.max_time is None else time.time() + self.max_time
        bar = self._progress()
        done = 0
        try:
            if self.rng_mode == "host":
                for i in range(self.proposals):
                    self.current_proposal = i
                    acc_i = self._host_proposal(i)
                    self._host_exchange(i)
                    if self.surrogate is not None and self.surrogate_fallback_proposal is None:
                        self._sur_recent.append(float(np.mean(acc_i)))
                        if len(self._sur_recent) >= self._sur_window:
                            self._surrogate_check(float(np.mean(self._sur_recent[-self._sur_window:])), i + 1)
                    if i % self.online_thinning == 0:
                        X, chi = self._chains.get_state()
                        self._write(X, chi)
                    done = i + 1
                    if bar is not None:
                        bar.update(1)
                        if i == 0 or i == self.proposals - 1 or i % 50 == 0:
                            bar.set_description(f"Tot. acc rate: {self._rate():.2f}. Progress", refresh=False)
                    if deadline is not None and time.time() > deadline:
                        raise TimeoutError
            else:
                chunk = self._chunk or max(self.online_thinning, min(self.proposals, 64))
                chunk = max(self.online_thinning, (chunk // self.online_thinning) * self.online_thinning)
                self._chains.set_trace(chunk // self.online_thinning + 1)
                # samples of chunk k are written to the files by a writer thread
                # while the device runs chunk k+1 (ctypes calls release the GIL)
                writer = _ChunkWriter(self._write)
                try:
                    while done < self.proposals:
                        n = min(chunk, self.proposals - done)
                        self._chains.reset_trace()
                        self._step, acc = self._chains.run(
                            n, self.amount_of_steps, self._step, self.integrator, self.randomize_stepsize,
                            self.autotuning, self.target_acceptance_rate, self.learning_rate,
                            self._device_seed, done, self.online_thinning)
                        self.accepted_proposals = self.accepted_proposals + (acc if B > 1 else int(acc[0]))
                        self._surrogate_check(float(np.sum(acc)) / (n * B), done + n)
                        cnt = self._chains.trace_count()
                        if cnt:
                            writer.put(self._chains.read_trace(0, cnt))
                        done += n
                        self.current_proposal = done - 1
                        if bar is not None:
                            bar.update(n)
                            bar.set_description(f"Tot. acc rate: {self._rate():.2f}. Progress", refresh=False)
                        if deadline is not None and time.time() > deadline:
                            raise TimeoutError
                finally:
                    writer.close()
        except KeyboardInterrupt:
            # Samplers.py:688-690: the interrupted proposal is dropped
            self.current_proposal = max(done - 1, 0)
        except TimeoutError:
            pass
        finally:
            if bar is not None:
                bar.close()

    def _rate(self):
        acc = np.sum(self.accepted_proposals) / self.chains
        return acc / (self.current_proposal + 1)

    def _host_proposal(self, i):
        B, d = self.chains, self.dimensions
        P0 = np.empty((B, d))
        eps = np.empty(B)
        u = np.empty(B)
        mm = self.mass_matrix
        for c, r in enumerate(self._rngs):
            P0[c] = mm.generate_momentum(rng=r).reshape(-1)
            eps[c] = r.uniform(0.5, 1.5) * self._step[c] if self.randomize_stepsize else self._step[c]
            # the accept uniform is drawn after the trajectory in the reference; no
            # other draw happens in between, so drawing it here consumes the same value
            u[c] = r.uniform(0, 1)
        res = self._chains.propose(P0, eps, u, self.amount_of_steps, self.integrator)
        with np.errstate(over="ignore", invalid="ignore"):
            a = np.exp(res["H0"] - res["H1"])
        if self.autotuning:
            for c in range(B):
                self.acceptance_rates[i, c] = a[c]
                self.stepsizes[i, c] = self._step[c]
                w = (i + 1) ** (-self.learning_rate)
                ac = 0 if np.isnan(a[c]) else a[c]
                s = self._step[c] - w * (self.target_acceptance_rate - min(ac, 1))
                if s <= 0:
                    s = max(s, self.minimal_stepsize)
                self._step[c] = s
        acc = res["accepted"]
        if B == 1:
            self.accepted_proposals += int(acc[0])
        else:
            self.accepted_proposals += acc.astype(np.int64)
        return acc

    # ------------------------------------------------------------------ #
    def _close_files(self):
        for w in self._writers:
            w.close()

    def _close(self):
        self.end_time = datetime.now()
        B, d = self.chains, self.dimensions
        self._chains_exchange_counts = self._chains.exchange_counts()
        X, chi = self._chains.get_state()
        self.current_model = X[0].reshape(d, 1) if B == 1 else X.T.copy()
        self.current_x = float(chi[0]) if B == 1 else chi
        self.stepsize = float(self._step[0]) if B == 1 else self._step.copy()
        if self.autotuning:
            if self.rng_mode == "device":
                rates, steps = self._chains.tuning_history()
                self.acceptance_rates, self.stepsizes = rates, steps
            # Samplers.py:1421-1422 keeps the entries before current_proposal
            self.acceptance_rates = self.acceptance_rates[: self.current_proposal]
            self.stepsizes = self.stepsizes[: self.current_proposal]
        rate = np.sum(self.accepted_proposals) / B / (self.current_proposal + 1)
        for c, w in enumerate(self._writers):
            acc_c = self.accepted_proposals if B == 1 else self.accepted_proposals[c]
            w.write_attribute("acceptance_rate", acc_c / (self.current_proposal + 1))
            w.write_attribute("end_time", self.end_time.strftime("%d-%b-%Y (%H:%M:%S.%f)"))
            w.write_attribute("runtime", str(self.end_time - self.start_time))
            w.write_attribute("runtime_seconds", (self.end_time - self.start_time).total_seconds())
            if self.autotuning:
                w.write_attribute("acceptance_rates", self.acceptance_rates[:, c:c + 1])
                w.write_attribute("stepsizes", self.stepsizes[:, c:c + 1])
        self._close_files()
        if self._memory is not None:
            self.results = (np.stack(self._memory, axis=2) if self._memory
                            else np.empty((B, d + 1, 0)))
        if self.diagnostic_mode:
            total = (self.end_time - self.start_time).total_seconds()
            names = ["residual GEMM / fused proposal", "endpoint misfit", "element-wise / Metropolis",
                     "gradient GEMM + kick/drift", "premultiplied stage GEMM + kick/drift",
                     "surrogate GEMMs", "surrogate kick/drift", "premultiplied split-tail epilogue"]
            self.diagnostics = {}
            print("Detailed statistics:")
            print(f"Total runtime: {total:.2f} seconds")
            print("{:<30} {:<12} {:<10}".format("Kernel group", "GPU ms", "launches"))
            for k, nm in enumerate(names):
                ms, n = self._chains.timing(k)
                self.diagnostics[nm] = (ms, n)
                print("{:<30} {:<12.3f} {:<10d}".format(nm, ms, n))
        self.acceptance_rate = rate

    def get_diagnostics(self):
        if not self.diagnostic_mode:
            raise AssertionError("Can't return diagnostics if sampler is not in diagnostic mode")
        return self.diagnostics

    def load_results(self, burn_in: int = 0) -> np.ndarray:
        """(d+1, N) array of chain 0 (or of the only chain)."""
        assert burn_in >= 0
        if self.samples_filename is None:
            return self.results[0][:, burn_in:]
        with Samples(self._filenames[0], burn_in=burn_in) as s:
            return np.array(s.numpy)


class _ChunkWriter:
    """Single background thread appending trace chunks (X (k, B, d), chi (k, B))
    to the samples files in order; close() drains the queue and re-raises a
    writer error in the caller."""

    def __init__(self, write):
        import queue
        import threading

        self._write = write
        self._q = queue.Queue(maxsize=4)
        self._err = None
        self._t = threading.Thread(target=self._run, daemon=True)
        self._t.start()

    def _run(self):
        while True:
            item = self._q.get()
            if item is None:
                return
            if self._err is not None:
                continue
            try:
                X, chi = item
                for k in range(X.shape[0]):
                    self._write(X[k], chi[k])
            except BaseException as e:   # reported by close()
                self._err = e

    def put(self, item):
        if self._err is not None:
            raise self._err
        self._q.put(item)

    def close(self):
        self._q.put(None)
        self._t.join()
        if self._err is not None:
            raise self._err


class ParallelSampleSMP:
    """Several chains with (possibly tempered) posteriors and replica exchange,
    mirroring ``hmclab.Samplers.ParallelSampleSMP`` (Samplers.py:1804-1977).

    The reference runs one process per chain and exchanges models through
    pipes; here all chains run as one batch on the GPU.  The posteriors must
    share one target up to their ``temperature`` (chain c samples
    exp(-chi / T_c)), which is how tempered ladders are built; the exchange
    schedule is drawn exactly as the reference draws it
    (``rng.choice(N, 2 * (N // 2), replace=False)`` per exchange event), and
    the second chain of each pair decides with a uniform from its own sampler's
    generator, so a seeded run reproduces the reference's decisions.
    Deviation: after an accepted swap the misfit moves with the model (the
    reference keeps the old ``current_x``, Samplers.py:640-668, so its samples
    file records a stale misfit and later exchange decisions use it).
    """

    def __init__(self, seed=None):
        self.rng = np.random.default_rng() if seed is None else np.random.default_rng(seed)

    def sample(self, samplers, filenames, posteriors, overwrite_existing_files=False, proposals: int = 100,
               exchange: bool = True, exchange_interval: int = 1, initial_model=None, kwargs=None):
        assert overwrite_existing_files, (
            "You have to manually enable overwriting samples. This is for safety. The "
            "existing file dialog doesn't work in the parallel case. "
            "Set `overwrite_existing_files=True`."
        )
        number_of_chains = len(samplers)
        assert len(filenames) == number_of_chains, (
            f"The number of supplied initial models ({len(filenames)}) "
            f"is not equal to the amount of chains ({number_of_chains}). "
            f"Supply {number_of_chains} models."
        )
        assert len(posteriors) == number_of_chains, (
            f"The number of supplied initial models ({len(posteriors)}) "
            f"is not equal to the amount of chains ({number_of_chains}). "
            f"Supply {number_of_chains} posteriors."
        )
        if type(initial_model) == list:
            assert len(initial_model) == number_of_chains, (
                f"The number of supplied initial models ({len(initial_model)}) "
                f"is not equal to the amount of chains ({number_of_chains}). "
                f"Supply either 1 or {number_of_chains} models."
            )
        if type(kwargs) == list:
            assert len(kwargs) == number_of_chains, (
                f"The number of supplied kwargs dictionaries ({len(kwargs)}) "
                f"is not equal to the amount of chains ({number_of_chains}). "
                f"Supply either 1 or {number_of_chains} kwargs. Note that even if you "
                f"don't want to pass arguments to one of the chains, and empty "
                f"dictionary still has to be included."
            )
        # Samplers.py:1880-1893: one row of distinct chain indices per exchange event
        schedule = None
        if exchange:
            exchanges_per_proposal = int(np.floor(number_of_chains / 2))
            schedule = np.vstack([
                self.rng.choice(number_of_chains, exchanges_per_proposal * 2, replace=False)
                for _ in range(int(proposals / exchange_interval))
            ])
        self.exchange_schedule = schedule
        _check_same_target(posteriors)
        temperatures = [float(getattr(p, "temperature", 1.0)) for p in posteriors]
        d = posteriors[0].dimensions
        if type(initial_model) == list:
            m0 = np.hstack([np.asarray(m, dtype=np.float64).reshape(d, 1) for m in initial_model])
        elif initial_model is None:
            m0 = None
        else:
            m0 = np.repeat(np.asarray(initial_model, dtype=np.float64).reshape(d, 1), number_of_chains, axis=1)
        kw = dict(kwargs[0] if type(kwargs) == list else (kwargs or {}))
        if type(kwargs) == list:
            for other in kwargs[1:]:
                assert other == kwargs[0], "all chains of one batch must share their sampler settings"
        for key in ("proposals", "overwrite_existing_file", "queue"):
            kw.pop(key, None)
        self.sampler = HMC(seed=getattr(samplers[0], "seed", None))
        names = None if all(f is None for f in filenames) else list(filenames)   # None: keep in memory
        self.sampler.sample(names, posteriors[0], proposals=proposals, initial_model=m0,
                            overwrite_existing_file=True, chains=number_of_chains, rng="host",
                            rngs=[s.rng for s in samplers], temperatures=temperatures,
                            exchange_schedule=schedule, exchange_interval=exchange_interval, **kw)
        self.exchange_counts = self.sampler._chains_exchange_counts
        return self


def _check_same_target(posteriors):
    """All posteriors of one batch must compile to the same device problem."""
    from .Distributions import _Spec

    def spec_of(p):
        s = _Spec()
        p._contribute(s)
        return s

    first = spec_of(posteriors[0])
    for p in posteriors[1:]:
        assert p.dimensions == posteriors[0].dimensions, "posteriors differ in dimension"
        s = spec_of(p)
        same = len(s.priors) == len(first.priors) and all(
            a[0] == b[0] and np.array_equal(np.asarray(a[1]), np.asarray(b[1]))
            and np.array_equal(np.asarray(a[2]), np.asarray(b[2])) for a, b in zip(s.priors, first.priors))
        if (s.likelihood is None) != (first.likelihood is None):
            same = False
        elif s.likelihood is not None and s.likelihood is not first.likelihood:
            ga, da, va = s.likelihood._whitened()
            gb, db, vb = first.likelihood._whitened()
            same = same and (ga.shape == gb.shape) and np.array_equal(np.asarray(da), np.asarray(db)) and (
                np.array_equal(ga.toarray() if hasattr(ga, "toarray") else ga,
                               gb.toarray() if hasattr(gb, "toarray") else gb))
            same = same and np.array_equal(np.asarray(va), np.asarray(vb))
        if not same:
            raise NotImplementedError(
                "the GPU batch needs posteriors that differ only by their temperature")


"""Cellular DMFT on the ED solver: the plain loop and its batched form.

Plain loop (one instance), the semantics of pyqcm's CDMFT class:
  * frequency grid and weights: ``frequency_grid`` (pyqcm/cdmft.py:821-899), default
    Legendre (1, 10, 5, 10, 5); distance weights are all 1;
  * one CDMFT step (pyqcm/cdmft.py:479-627): solve the impurity problem at the current
    bath parameters x, build the host G_host (src_qcm/CPT.cpp:356-424), then minimize the
    distance (src_qcm/CPT.cpp:519-555) over the bath parameters with adaptive Nelder-Mead
    from an initial simplex of step ``initial_step`` (pyqcm/cdmft.py:1614-1633; xatol =
    accur_bath, fatol = accur_dist);
  * outer iteration: Broyden on F(x) = x - step(x) with xtol 1e-6 (pyqcm/__init__.py:2844-2898);
  * convergence: ``convergence_manager`` on 'parameters' (||x - x_i|| / len(x) below
    ``accur`` against the last ``depth`` iterates, pyqcm/cdmft.py:14-106).

Bath fit: for a diagonal bath whose parameters act linearly (the models here), all the
fits of an outer iteration run in one device launch (qed_cdmft_fit_batch, one wavefront
per instance). The same algorithm also exists as a host coroutine whose trial points are
evaluated on the device (qed_cdmft_distance_multi); it reproduces scipy's Nelder-Mead
call for call and is the form used for general baths.

Batched loop: B independent CDMFT problems advance in lockstep; the ED solves of one
outer iteration and their bath fits go to the device together.
"""
from __future__ import annotations

import time

import numpy as np

from . import lattice as LT


class MinimizationError(RuntimeError):
    pass


class TooManyIterationsError(RuntimeError):
    pass


# ----------------------------------------------------------------- frequency grid

class frequency_grid:
    """Imaginary-frequency grid of the distance function (pyqcm/cdmft.py:821-899):
    'legendre' with specs (w1, w2, n1, n2, n3), 'matsubara' (wc, beta), 'regular'
    (wc, n1, n2). cdmft_weight is 1 on every frequency."""

    def __init__(self, grid_type="legendre", specs=(1, 10, 5, 10, 5)):
        if grid_type.lower() == "legendre":
            w1, w2, n1, n2, n3 = specs
            self.wr, self.weight = LT.legendre_frequency_grid(w1, w2, n1, n2, n3)
            self.name = "legendre({:g}-{:g}-{:d}-{:d}-{:d})".format(w1, w2, n1, n2, n3)
        elif grid_type.lower() == "matsubara":
            wc, beta = specs
            self.wr = np.arange(np.pi / beta, wc + 1e-6, 2 * np.pi / beta)
            self.weight = (2 * np.pi / beta) * np.ones_like(self.wr)
            self.name = "matsubara({:g}-{:g})".format(wc, beta)
        elif grid_type == "regular":
            wc, n1, n2 = specs
            self.wr, self.weight = regular_frequency_grid(wc, n1, n2)
            self.name = "regular({:g}-{:d}-{:d})".format(wc, n1, n2)
        else:
            raise ValueError("Unknown type of frequency grid " + grid_type)
        self.nw = len(self.wr)
        self.cdmft_weight = np.ones_like(self.wr)


def regular_frequency_grid(wc, n1, n2):
    """n1 evenly spaced points in (0, wc) and n2 Gauss-Legendre points on [wc, inf) in
    u = 1/w (weights with 1/pi)."""
    dw = wc / n1
    w = np.concatenate([(np.arange(n1) + 0.5) * dw, np.zeros(n2)])
    p = np.concatenate([np.full(n1, dw), np.zeros(n2)])
    x, q = np.polynomial.legendre.leggauss(n2)
    u = 0.5 * (1.0 / wc) * (x[::-1] + 1.0)
    w[n1:] = 1.0 / u
    p[n1:] = 0.5 * (1.0 / wc) * q[::-1] * w[n1:] ** 2
    return w, p / np.pi


# ------------------------------------------------------------ convergence, Broyden

class convergence_manager:
    """Convergence on a quantity compared with its `depth` predecessors
    (pyqcm/cdmft.py:14-106)."""

    def __init__(self, name, diff_func, tol, depth=2):
        self.name, self.diff_func, self.tol, self.depth = name, diff_func, tol, depth
        self.diff = np.zeros(depth)
        self.iter = 0
        self.X = []
        self.converged = False

    def test(self, x):
        self.iter += 1
        if self.iter <= self.depth:
            self.X.append(x)
            return False
        for i in range(self.depth):
            self.diff[i] = self.diff_func(x, self.X[i])
        converged = bool(np.all(self.diff[: self.depth] <= self.tol))
        for i in range(self.depth - 1):
            self.X[i] = self.X[i + 1]
        self.X[-1] = x
        self.converged = converged
        return converged


def parameters_convergence(tol, depth=2):
    return convergence_manager("parameters", lambda x, y: np.linalg.norm(x - y) / x.shape[0],
                               tol, depth)


def broyden(F, x0, iJ0=0.0, xtol=1e-6, convergence_test=None, maxiter=32, miniter=0):
    """Broyden root search with the good inverse-Jacobian update and custom convergence
    tests (pyqcm/__init__.py:2844-2898). Returns (x, iterations, inverse Jacobian)."""
    n = len(x0)
    I = (1.0 + iJ0) * np.eye(n) if np.isscalar(iJ0) else np.array(iJ0, float)
    f0 = F(x0)
    f = np.copy(f0)
    x = np.array(x0, float)
    it = 0
    while True:
        dx_vec = -I @ f
        if np.any(np.isnan(dx_vec)):
            raise ValueError("NaN produced in Broyden method")
        x = x + dx_vec
        if np.linalg.norm(dx_vec) < xtol and it >= miniter:
            if convergence_test is not None:
                convergence_test()
            break
        if convergence_test is not None and convergence_test() and it >= miniter:
            break
        f = F(x)
        df = f - f0
        f0 = np.copy(f)
        T = np.outer(dx_vec - I @ df, dx_vec)
        I = I + T @ I / (dx_vec @ (I @ df))
        it += 1
        if it > maxiter:
            raise TooManyIterationsError(maxiter)
    return x, it, I


# ------------------------------------------------------------------- Nelder-Mead

def nelder_mead_coroutine(x0, initial_step, xatol, fatol, maxfev, stable_sort=False):
    """Adaptive Nelder-Mead as scipy.optimize.minimize(method='Nelder-Mead',
    options={'adaptive': True, 'initial_simplex': x0 + step e_i, 'xatol', 'fatol',
    'maxfev'}) performs it (the options used at pyqcm/cdmft.py:1614-1633).

    A generator: it yields 2-D arrays of points to evaluate and receives their values;
    it returns (x, nfev, success, fun). With the default sort it reproduces scipy's
    sequence of evaluations exactly. scipy orders the simplex with numpy's default
    argsort, whose order for exactly equal values depends on the sort implementation;
    stable_sort=True keeps equal values in their previous order, the rule the device
    fit (qed_cdmft_fit_batch) applies."""
    x0 = np.asarray(x0, float)
    N = len(x0)
    rho, chi, psi, sigma = 1.0, 1.0 + 2.0 / N, 0.75 - 1.0 / (2.0 * N), 1.0 - 1.0 / N
    kind = "stable" if stable_sort else None
    sim = np.repeat(x0[None], N + 1, 0)
    for i in range(N):
        sim[i + 1, i] += initial_step
    nfev = 0
    fsim = np.asarray((yield sim.copy()), float).copy()
    nfev += N + 1
    ind = np.argsort(fsim, kind=kind)
    sim, fsim = sim[ind], fsim[ind]
    while nfev < maxfev:
        if (np.max(np.abs(sim[1:] - sim[0])) <= xatol and np.max(np.abs(fsim[0] - fsim[1:])) <= fatol):
            break
        xbar = np.add.reduce(sim[:-1], 0) / N
        xr = (1 + rho) * xbar - rho * sim[-1]
        fxr = float((yield xr[None])[0])
        nfev += 1
        doshrink = False
        if fxr < fsim[0]:
            xe = (1 + rho * chi) * xbar - rho * chi * sim[-1]
            fxe = float((yield xe[None])[0])
            nfev += 1
            if fxe < fxr:
                sim[-1], fsim[-1] = xe, fxe
            else:
                sim[-1], fsim[-1] = xr, fxr
        elif fxr < fsim[-2]:
            sim[-1], fsim[-1] = xr, fxr
        else:
            if fxr < fsim[-1]:
                xc = (1 + psi * rho) * xbar - psi * rho * sim[-1]
                fxc = float((yield xc[None])[0])
                nfev += 1
                if fxc <= fxr:
                    sim[-1], fsim[-1] = xc, fxc
                else:
                    doshrink = True
            else:
                xcc = (1 - psi) * xbar + psi * sim[-1]
                fxcc = float((yield xcc[None])[0])
                nfev += 1
                if fxcc < fsim[-1]:
                    sim[-1], fsim[-1] = xcc, fxcc
                else:
                    doshrink = True
            if doshrink:
                pts = sim[0] + sigma * (sim[1:] - sim[0])
                vals = np.asarray((yield pts), float)
                nfev += N
                sim[1:] = pts
                fsim[1:] = vals
        ind = np.argsort(fsim, kind=kind)
        sim, fsim = sim[ind], fsim[ind]
    return sim[0].copy(), nfev, nfev < maxfev, float(np.min(fsim))


def run_coroutines(coros, evaluate):
    """Drives optimizer coroutines in lockstep. evaluate(list of (index, points)) returns
    the list of value arrays; returns the list of coroutine results."""
    n = len(coros)
    results = [None] * n
    pending = {}
    for i, c in enumerate(coros):
        pending[i] = next(c)
    while pending:
        order = list(pending.keys())
        vals = evaluate([(i, pending[i]) for i in order])
        nxt = {}
        for i, v in zip(order, vals):
            try:
                nxt[i] = coros[i].send(v)
            except StopIteration as stop:
                results[i] = stop.value
        pending = nxt
    return results


# ------------------------------------------------------------------ the problem

class CDMFTProblem:
    """A cluster-plus-bath model on a lattice with a set of variational bath parameters.

    `model` is a lattice.LatticeModel whose bath operators were added with
    add_cluster_operator; `varia` names the bath parameters (optionally with the
    cluster suffix '_1', as in pyqcm); `sectors` is the target-sector string, or a function
    of the parameter values returning it."""

    def __init__(self, model, varia, sectors, grid=None, kside=32, graph=True, nambu=None):
        self.model = model
        # the spinor of the fit: None until the first solve tells the instance's mixing
        # (1: pairing terms, the Nambu spinor; 2: spin flip, the spinor (c_up, c_down);
        # 4: the two spin blocks side by side; 0: one spin block); nambu is true for every
        # two-block form, spinor_mixing says which
        self.nambu = nambu
        self.spinor_mixing = 1 if nambu else 0
        self.graph = graph  # replay the batched Lanczos steps as hipGraphs
        self.varia = [v[:-2] if v.endswith("_1") else v for v in varia]
        self.sectors = sectors
        self.grid = grid or frequency_grid()
        self.kside = kside

    def solve(self, values, x, label=None):
        vals = dict(values)
        vals.update(zip(self.varia, np.asarray(x, float)))
        inst = self.model.instance(vals, self.sectors_of(values), label)
        inst.ground_state()
        self._check_mixing(inst)
        return inst

    def _check_mixing(self, inst):
        m = inst.mixing
        if m not in (0, 1, 2, 4):
            raise RuntimeError("CDMFT takes the normal, Nambu, spin-flip and up/down mixings "
                               "(instance mixing %d)" % m)
        if self.nambu is None:
            self.nambu = m in (1, 2, 4)
            self.spinor_mixing = m if m in (1, 2, 4) else 0
        elif self.nambu != (m in (1, 2, 4)) or (self.nambu and self.spinor_mixing != m):
            raise RuntimeError("the instance's mixing changed during CDMFT (%d)" % m)

    def host(self, inst):
        """G_host; for the up/down mixing the two spin blocks side by side (2n x 2n)."""
        if inst.mixing == 4:
            up = inst.cdmft_host(self.grid.wr, self.kside, 0)
            dn = inst.cdmft_host(self.grid.wr, self.kside, 1)
            n = up.shape[1]
            out = np.zeros((up.shape[0], 2 * n, 2 * n), complex)
            out[:, :n, :n], out[:, n:, n:] = up, dn
            return out
        return inst.cdmft_host(self.grid.wr, self.kside)

    def solve_batch(self, values_list, x_list, graph=None):
        """Host functions of many instances: one batched device solve for all of them
        (qcm.batch_solve), then the k-sums per instance."""
        from . import qcm
        insts = []
        for v, x in zip(values_list, x_list):
            vals = dict(v)
            vals.update(zip(self.varia, np.asarray(x, float)))
            insts.append(self.model.instance(vals, self.sectors_of(v)))
        self.last_stats = qcm.batch_solve([I.label for I in insts], graph=self.graph if graph is None else graph)
        hosts = []
        for I in insts:
            E0 = I.ground_state()[0]
            if np.isfinite(E0):
                self._check_mixing(I)
                hosts.append(self.host(I))
            else:  # the impurity solve failed for this instance
                n = self.model.n * (2 if self.nambu else 1)
                hosts.append(np.full((len(self.grid.wr), n, n), np.nan, complex))
            I.erase()
        return hosts

    def sectors_of(self, values):
        return self.sectors(values) if callable(self.sectors) else self.sectors

    def distance(self, Ghost, X):
        eb, tcb = self.model.diagonal_bath(self.varia, X, mixing=self.spinor_mixing)
        return LT.cdmft_distance_batch(self.grid.wr, self.grid.cdmft_weight, Ghost, eb, tcb,
                                       self.spinor_mixing)

    def distance_multi(self, hosts, requests):
        """requests: list of (host index, points (m x nvar)); all points are evaluated in
        one device call (qed_cdmft_distance_multi), values returned per request."""
        pts = np.concatenate([np.atleast_2d(p) for _, p in requests])
        host_of = np.concatenate([np.full(len(np.atleast_2d(p)), h, np.int32) for h, p in requests])
        eb, tcb = self.model.diagonal_bath(self.varia, pts, mixing=self.spinor_mixing)
        d = LT.cdmft_distance_multi(self.grid.wr, self.grid.cdmft_weight, np.asarray(hosts),
                                    host_of, eb, tcb, self.spinor_mixing)
        out, o = [], 0
        for _, p in requests:
            m = len(np.atleast_2d(p))
            out.append(d[o:o + m])
            o += m
        return out

    def linear_bath_map(self):
        """(E (nvar x nb), T (nvar x n x nb)) when every variational parameter acts on a
        diagonal bath (eb and tcb linear in x) of one spin block; None otherwise."""
        if self.nambu:
            return None
        nb, n = self.model.n_bath, self.model.n
        E = np.zeros((len(self.varia), nb))
        T = np.zeros((len(self.varia), n, nb))
        for p, nm in enumerate(self.varia):
            Bm, Cm = self.model.bath_ops[nm]
            if np.any(Bm - np.diag(np.diag(Bm))):
                return None
            E[p] = np.diag(Bm)
            T[p] = Cm
        return E, T

    def fit(self, hosts, host_of, X0, initial_step=0.1, xatol=1e-6, fatol=1e-8, maxfev=1_000_000,
            where="device"):
        """Nelder-Mead bath fits of len(X0) instances; instance b fits to hosts[host_of[b]].
        'device': one launch of qed_cdmft_fit_batch; 'host': the coroutine form of the same
        algorithm, its trial points evaluated on the device. Returns (x, f, nfev, ok)."""
        X0 = np.atleast_2d(np.asarray(X0, float))
        hosts = np.asarray(hosts)
        host_of = np.ascontiguousarray(host_of, np.int32)
        lin = self.linear_bath_map() if where == "device" else None
        if lin is not None:
            return LT.cdmft_fit_batch(self.grid.wr, self.grid.cdmft_weight, hosts, host_of, lin[0],
                                      lin[1], X0, initial_step, xatol, fatol, maxfev)
        coros = [nelder_mead_coroutine(x, initial_step, xatol, fatol, maxfev, stable_sort=True)
                 for x in X0]
        res = run_coroutines(coros, lambda reqs: self.distance_multi(
            hosts, [(host_of[i], p) for i, p in reqs]))
        x = np.array([r[0] for r in res])
        return x, np.array([r[3] for r in res]), np.array([r[1] for r in res]), \
            np.array([r[2] for r in res])


class CDMFT:
    """Plain CDMFT loop for one instance (pyqcm/cdmft.py:194-478), Broyden iteration,
    Nelder-Mead bath fit, convergence on 'parameters'."""

    def __init__(self, problem, values, x0, maxiter=32, miniter=0, depth=2, accur=1e-4,
                 accur_bath=1e-6, accur_dist=1e-8, initial_step=0.1,
                 max_function_eval=1_000_000, iteration="broyden", label=None, verbose=False,
                 fit="device"):
        self.problem = problem
        self.fit_where = fit
        self.values = dict(values)
        self.x = np.array(x0, float)
        self.accur_bath, self.accur_dist = accur_bath, accur_dist
        self.initial_step = initial_step
        self.maxfev = max_function_eval
        self.label = label
        self.verbose = verbose
        self.conv = parameters_convergence(accur, depth)
        self.n_solves = 0
        self.nfev = 0
        self.dist = None
        self.history = []
        self.t_ed = self.t_fit = 0.0

        def F(x):
            self.x = np.copy(x)
            self.step()
            return x - self.x

        def G():
            return self.conv.test(np.copy(self.x))

        if iteration == "broyden":
            self.x, self.niter, _ = broyden(F, self.x, 0.0, xtol=1e-6, convergence_test=G,
                                            maxiter=maxiter, miniter=miniter)
        elif iteration == "fixed_point":
            it = 0
            while True:
                F(self.x)
                if G() and it >= miniter:
                    break
                it += 1
                if it > maxiter:
                    raise TooManyIterationsError(maxiter)
            self.niter = it
        else:
            raise ValueError("iteration must be 'broyden' or 'fixed_point'")

    def step(self):
        t0 = time.perf_counter()
        inst = self.problem.solve(self.values, self.x, self.label)
        Gh = self.problem.host(inst)
        E0 = inst.ground_state()[0]
        inst.erase()
        self.n_solves += 1
        t1 = time.perf_counter()
        X, F, NF, OK = self.problem.fit(Gh[None], [0], self.x[None], self.initial_step,
                                        self.accur_bath, self.accur_dist, self.maxfev, self.fit_where)
        xo, fun, nfev, ok = X[0], float(F[0]), int(NF[0]), bool(OK[0])
        if not ok:
            raise MinimizationError("Failure of the optimization of the bath parameters")
        if np.any(np.isnan(xo)):
            raise MinimizationError("NaN found in optimization of bath parameters")
        self.nfev += nfev
        self.dist = fun
        self.x = xo
        self.history.append((np.copy(xo), fun, E0))
        self.t_ed += t1 - t0
        self.t_fit += time.perf_counter() - t1
        if self.verbose:
            print("CDMFT step %d: distance %.4g, %d evaluations" % (len(self.history), fun, nfev),
                  flush=True)


class BatchedCDMFT:
    """Many independent CDMFT problems (same model, different parameter values and
    starting points), each following the plain loop's semantics. Up to `capacity`
    problems are active at a time; each round, the impurity solves and bath fits of all
    active problems go to the device together, and a problem that converges or fails
    leaves the batch and the next queued problem takes its place, so the batch stays full
    until the queue is empty.

    `solver(values_list, x_list)` returns the host functions of the listed instances;
    "batched" solves them in one device call (qcm.batch_solve); the default solves them
    one after the other through `problem`."""

    def __init__(self, problem, values_list, x0_list, maxiter=32, depth=2, accur=1e-4,
                 accur_bath=1e-6, accur_dist=1e-8, initial_step=0.1, max_function_eval=1_000_000,
                 solver=None, verbose=False, fit="device", capacity=None, alpha_list=None,
                 fallback_list=None, fallback_after=None):
        """alpha_list: per problem, the Broyden inverse-Jacobian start (the reference's CDMFT
        `alpha`: a float gives (1 + alpha) times the unit matrix, a matrix is used as the
        inverse Jacobian, e.g. from a previous run; pyqcm/cdmft.py:171, 414-421). The
        inverse Jacobians reached are kept in self.alpha.

        fallback_list: per problem, a second starting point (e.g. the cold start when x0 is
        a predicted warm start). A problem whose bath fit fails, whose Broyden step is
        undefined, or that has not converged after `fallback_after` iterations (default:
        maxiter) starts again from it once, with the unit inverse Jacobian and a fresh
        convergence test; its solves and iterations add up. self.restarted marks them."""
        self.problem = problem
        self.alpha_in = alpha_list
        self.fallback = None if fallback_list is None else [np.array(x, float) for x in fallback_list]
        self.fallback_after = maxiter if fallback_after is None else int(fallback_after)
        self.fit_where = fit
        self.capacity = capacity or len(values_list)
        B = len(values_list)
        self.values = [dict(v) for v in values_list]
        self.x = [np.array(x, float) for x in x0_list]
        self.accur_bath, self.accur_dist = accur_bath, accur_dist
        self.initial_step, self.maxfev = initial_step, max_function_eval
        self.maxiter = maxiter
        if solver == "batched":
            solver = problem.solve_batch
        self.solver = solver or self._solve_sequential
        self.verbose = verbose
        self.conv = [parameters_convergence(accur, depth) for _ in range(B)]
        self._accur, self._depth = accur, depth
        self.restarted = np.zeros(B, bool)
        self.n_solves = np.zeros(B, int)
        self.nfev = np.zeros(B, int)
        self.dist = np.full(B, np.nan)
        self.niter = np.zeros(B, int)
        self.converged = np.zeros(B, bool)
        self.failed = np.zeros(B, bool)
        self.t_ed = self.t_fit = 0.0
        self._run()

    def _solve_sequential(self, values_list, x_list):
        hosts = []
        for v, x in zip(values_list, x_list):
            inst = self.problem.solve(v, x)
            hosts.append(self.problem.host(inst))
            inst.erase()
        return hosts

    def _steps(self, idx, xs):
        """One CDMFT step for the instances idx at bath parameters xs: returns new xs."""
        t0 = time.perf_counter()
        hosts = self.solver([self.values[i] for i in idx], xs)
        t1 = time.perf_counter()
        hosts = np.asarray(hosts)
        good = [j for j in range(len(idx)) if np.all(np.isfinite(hosts[j]))]
        X = np.array(xs, float)
        F = np.full(len(idx), np.nan)
        NF = np.zeros(len(idx), int)
        OK = np.zeros(len(idx), bool)
        if good:
            Xg, Fg, NFg, OKg = self.problem.fit(hosts[good], np.arange(len(good)), X[good],
                                                self.initial_step, self.accur_bath, self.accur_dist,
                                                self.maxfev, self.fit_where)
            X[good], F[good], NF[good], OK[good] = Xg, Fg, NFg, OKg
        self.t_ed += t1 - t0
        self.t_fit += time.perf_counter() - t1
        out = []
        for j, i in enumerate(idx):
            xo, fun, nfev, ok = X[j], float(F[j]), int(NF[j]), bool(OK[j])
            self.n_solves[i] += 1
            self.nfev[i] += nfev
            self.dist[i] = fun
            if not ok or np.any(np.isnan(xo)):
                self.failed[i] = True
            out.append(xo)
        return out

    def _run(self):
        """Broyden per problem (pyqcm/__init__.py:2844-2898) as a state machine: each round
        evaluates F at the pending point of every active problem (the start point first),
        then applies the inverse-Jacobian update, the step and the convergence tests."""
        from collections import deque
        B = len(self.x)
        n = len(self.x[0])
        Ij = []
        for i in range(B):
            a = 0.0 if self.alpha_in is None else self.alpha_in[i]
            if isinstance(a, tuple):
                raise TypeError("alpha must be a float or an inverse-Jacobian matrix")
            if np.isscalar(a):
                Ij.append((1.0 + a) * np.eye(n))
                continue
            a = np.array(a, float, copy=True)
            if a.shape != (n, n):
                raise ValueError("the initial Jacobian has the wrong dimensions")
            Ij.append(a)
        self.alpha = Ij
        f0 = [None] * B
        dxs = [None] * B
        it = np.zeros(B, int)
        it_done = np.zeros(B, int)  # iterations of abandoned first attempts
        started = np.zeros(B, bool)

        def restart(i):
            """the problem's second attempt from its fallback start; False when none is left"""
            if self.fallback is None or self.restarted[i]:
                return False
            self.restarted[i] = True
            self.failed[i] = False
            self.x[i] = np.copy(self.fallback[i])
            Ij[i] = np.eye(n)
            f0[i] = dxs[i] = None
            started[i] = False
            it_done[i] += it[i]
            it[i] = 0
            self.conv[i] = parameters_convergence(self._accur, self._depth)
            return True
        queue = deque(range(B))
        active = []
        self.rounds = 0
        while active or queue:
            while queue and len(active) < self.capacity:
                active.append(queue.popleft())
            pts = [np.copy(self.x[i]) for i in active]
            new = self._steps(active, pts)
            self.rounds += 1
            nxt = []
            for i, x, xn in zip(active, pts, new):
                if self.failed[i]:
                    if restart(i):
                        nxt.append(i)
                    continue
                fi = x - xn
                if not started[i]:
                    started[i] = True
                else:
                    df = fi - f0[i]
                    dx = dxs[i]
                    T = np.outer(dx - Ij[i] @ df, dx)
                    Ij[i] = Ij[i] + T @ Ij[i] / (dx @ (Ij[i] @ df))
                    it[i] += 1
                    if it[i] > self.maxiter or (it[i] > self.fallback_after and not self.restarted[i]
                                                and self.fallback is not None):
                        if restart(i):
                            nxt.append(i)
                        else:
                            self.failed[i] = True
                        continue
                f0[i] = fi.copy()
                dx = -Ij[i] @ fi
                if np.any(np.isnan(dx)):
                    if restart(i):
                        nxt.append(i)
                    else:
                        self.failed[i] = True
                    continue
                dxs[i] = dx
                self.x[i] = x + dx
                if np.linalg.norm(dx) < 1e-6:
                    self.conv[i].test(np.copy(self.x[i]))
                    self.converged[i] = True
                    continue
                if self.conv[i].test(np.copy(self.x[i])):
                    self.converged[i] = True
                    continue
                nxt.append(i)
            active = nxt
            if self.verbose:
                print("batched CDMFT: %d active, %d queued, %d converged, %d failed"
                      % (len(active), len(queue), self.converged.sum(), self.failed.sum()), flush=True)
        self.niter = it + it_done


# ------------------------------------------------------------- 2x2 + 8 bath model

PLAQUETTE_SEQ = [[1, 1, 1, 1], [1, 1, 1, 1], [1, 1, -1, -1], [1, 1, -1, -1],
                 [1, -1, 1, -1], [1, -1, 1, -1], [1, -1, -1, 1], [1, -1, -1, 1]]


def plaquette_8b_model(name="clus_2x2_8b"):
    """2x2 plaquette on the square lattice (t, tp, mu, U) with 8 bath orbitals: eb_x
    energies and tb_x site-bath hybridizations with the C2v sign patterns of the
    reference's tests/test_diagonalization/cluster_2x2_2C_8b_C2v_L.py:8-33 (trivial
    point group here, as the factorized solver requires)."""
    T = LT.square_tiling(2, 2)
    M = LT.LatticeModel(name, T, n_bath=8)
    L = M.L
    for x in range(1, 9):
        b = 4 + x
        M.add_cluster_operator("eb%d" % x, [(b, b, 1.0), (b + L, b + L, 1.0)])
    for x in range(1, 9):
        b = 4 + x
        seq = PLAQUETTE_SEQ[x - 1]
        el = [(i + 1, b, float(seq[i])) for i in range(4)] + \
             [(i + 1 + L, b + L, float(seq[i])) for i in range(4)]
        M.add_cluster_operator("tb%d" % x, el)
    return M


PLAQUETTE_VARIA = ["eb%d" % x for x in range(1, 9)] + ["tb%d" % x for x in range(1, 9)]


def plaquette_sectors(values):
    """Target sectors of the 2x2+8-bath cluster (12 orbitals) for chemical potential mu in
    [0, U]: the even particle number N0 = 12 + 2 round(4 (mu/U - 1/2)) (8 .. 16) and its
    even neighbours within 8 .. 16, all at S = 0."""
    U, mu = float(values["U"]), float(values["mu"])
    f = mu / U if U > 0 else 0.5
    n0 = 12 + 2 * int(round(4 * (min(max(f, 0.0), 1.0) - 0.5)))
    ns = [n for n in (n0 - 2, n0, n0 + 2) if 8 <= n <= 16]
    return "/".join("R0:N%d:S0" % n for n in ns)


def random_plaquette_instances(n, seed):
    """Seeded (model values, bath parameters) pairs: U ~ U[0.5, 12], mu ~ U[0, U],
    t = 1, t' ~ U[-0.5, 0.5], bath energies ~ U[-3, 3], hybridizations ~ U[0.1, 1]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        U = rng.uniform(0.5, 12.0)
        vals = {"t": 1.0, "tp": rng.uniform(-0.5, 0.5), "U": U, "mu": rng.uniform(0.0, U)}
        x = np.concatenate([rng.uniform(-3, 3, 8), rng.uniform(0.1, 1.0, 8)])
        out.append((vals, x))
    return out


def canonical_map(x):
    """The signed permutation P with canonical_bath(x) = P x for one 16-parameter point
    (sign flips of the tb, then the swaps of each pair); a Broyden inverse Jacobian J
    reached at x reads P J P^T in the canonical coordinates."""
    x = np.asarray(x, float)
    n = len(x)
    S = np.eye(n)
    for k in range(8, 16):
        if x[k] < 0:
            S[k, k] = -1.0
    perm = np.arange(n)
    for p in range(4):
        a, b = 2 * p, 2 * p + 1
        if x[a] > x[b]:
            perm[[a, b]] = perm[[b, a]]
            perm[[8 + a, 8 + b]] = perm[[8 + b, 8 + a]]
    Pm = np.eye(n)[perm]
    return Pm @ S


def canonical_bath(x):
    """Gauge-fixes 2x2+8-bath parameters: Gamma is invariant under tb_x -> -tb_x and
    under exchanging the two bath orbitals of one sign pattern; take tb >= 0 and order
    each pair by bath energy."""
    x = np.array(x, float)
    eb, tb = x[..., :8].copy(), np.abs(x[..., 8:])
    for p in range(4):
        a, b = 2 * p, 2 * p + 1
        swap = eb[..., a] > eb[..., b]
        ea, eb_ = eb[..., a].copy(), eb[..., b].copy()
        ta, tb_ = tb[..., a].copy(), tb[..., b].copy()
        eb[..., a] = np.where(swap, eb_, ea)
        eb[..., b] = np.where(swap, ea, eb_)
        tb[..., a] = np.where(swap, tb_, ta)
        tb[..., b] = np.where(swap, ta, tb_)
    return np.concatenate([eb, tb], -1)

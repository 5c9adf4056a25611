"""CPU restatement of the reference's thermal-balance outer loop
(3DPDR.F90:584-1241, built with THERMALBALANCE, GUESS_TEMP, TEMP_FIX,
CO_FIX and, optionally, PSEUDO_1D), one point at a time, on the oracle's C
routines (oracle/thermal.c: FIND_CCOEFF, escape_probability, solvlevpop,
LTE populations, dust temperature; oracle/heating.c: CALC_HEATING;
oracle/rates.c: CALCULATE_REACTION_RATES).

TEST INFRASTRUCTURE ONLY: the checker of the device loop
(pdrchem_cloud_thermal_iterate).  The chemistry of the loop (three
iterations of rates, solve and columns, 3DPDR.F90:602-636) is the device's,
checked against the oracle's solver by its own tests: `iterate` takes the
abundances and columns the chemistry left, and restates everything else.

State per point, named after the reference's arrays: T (gastemperature),
Tprev (previousgastemperature), Tlow, Thigh, prev (previouschange: '', 'H',
'C'), binarychop (dobinarychop), converged, doleveltmin, level_converged;
pop (pdr(p)%X_pop), sol (X_solution), cool (X_cool), heat (all_heating).
"""
from __future__ import annotations

import numpy as np

import oracle as O

THERMAL_SPECIES = ("NCx", "NC", "NO", "NCO", "NH", "NPROTON", "NELECT", "NHE", "NH2")


def line_slot(ci, il, jl):
    """index of the (upper il, lower jl) line of coolant ci among the 17 lines
    of the reference's .line.fin (3DPDR.F90:1407-1416), or -1"""
    if ci == 0:
        return 0 if (il, jl) == (1, 0) else -1
    if ci in (1, 2):
        b = 1 if ci == 1 else 4
        return {(1, 0): b, (2, 0): b + 1, (2, 1): b + 2}.get((il, jl), -1)
    return 7 + jl if (il == jl + 1 and jl < 10) else -1


class ThermalLoop:
    def __init__(self, name, net, coolants, rho, uv, G, Av, T0, count, offs, proj, step, *, Tmin=2.7, Tmax=30000.0,
                 Fcrit=0.005, Tdiff=0.01, pseudo_ray=-1, co_fix=True, temp_fix=True, dust=True, Td0=20.0,
                 phys=None):
        self.name, self.net = name, net
        self.tabs = [O.CoolantTables(c) for c in coolants]
        self.nlev = [c.nlev for c in coolants]
        self.off = np.concatenate([[0], np.cumsum(self.nlev)]).astype(int)
        self.nlevtot = int(self.off[-1])
        self.slots = [net.labels[k] - 1 for k in THERMAL_SPECIES]
        self.orc = O.Oracle(name)
        self.phys = phys or O.default_phys()
        self.rho, self.uv, self.G, self.Av = (np.asarray(a, dtype=np.float64) for a in (rho, uv, G, Av))
        self.count, self.offs, self.proj, self.step = count.reshape(-1), offs, proj, step
        self.n = n = len(self.rho)
        self.nrays = self.G.shape[1]
        self.Tmin, self.Tmax, self.Fcrit, self.Tdiff = Tmin, Tmax, Fcrit, Tdiff
        self.pseudo_ray, self.co_fix, self.temp_fix = pseudo_ray, co_fix, temp_fix
        # dust temperatures (3DPDR.F90:430-435) and the GUESS_TEMP brackets
        # (3DPDR.F90:1702-1716)
        self.Td = (np.array([O.dust_temperature(self.uv[p], self.G[p]) for p in range(n)]) if dust
                   else np.full(n, float(Td0)))
        self.T = np.array(T0, dtype=np.float64)
        self.Tprev = self.T.copy()
        self.Tlow = np.maximum(self.T / 2.0, Tmin)
        self.Thigh = np.minimum(self.T * 1.5, Tmax)
        self.prev = [""] * n
        self.binarychop = np.zeros(n, bool)
        self.converged = np.zeros(n, bool)
        self.doleveltmin = np.zeros(n, bool)
        self.level_converged = np.zeros(n, bool)
        self.pop = np.zeros((n, self.nlevtot))
        self.sol = np.zeros((n, self.nlevtot))
        self.cool = np.zeros((n, 4))
        self.total = np.zeros(n)
        self.lines = np.zeros((n, 17))
        self.tau_lines = np.zeros((n, 17))     # optical depths along tau_ray (.opdp.fin)
        self.tau_ray = pseudo_ray if pseudo_ray >= 0 else 6
        self.heat = np.zeros((n, 12))
        self.iteration = 0
        self.levpop_iteration = 0
        self.level_conv = False
        self.first_time = True
        self.pct = [0, 0, 0, 0]
        self.started = False
        self.done = False

    # -- level populations --------------------------------------------------
    def lte(self, ab, mask):
        """CALCULATE_LTE_POPULATIONS of every coolant (3DPDR.F90:520-552)"""
        for p in np.where(mask)[0]:
            for ci, tab in enumerate(self.tabs):
                dens = ab[p, self.slots[ci]] * self.rho[p]
                self.pop[p, self.off[ci]:self.off[ci + 1]] = O.lte_populations(tab, self.T[p], dens)

    def chains(self, p, ci):
        """per ray (populations of the point and its evaluation points, steps)"""
        a, b = self.off[ci], self.off[ci + 1]
        out = []
        for j in range(self.nrays):
            o, m = self.offs[p * self.nrays + j], self.count[p * self.nrays + j]
            rows = np.concatenate([[p], self.proj[o:o + m]]).astype(int)
            out.append((self.pop[rows, a:b], self.step[o:o + m]))
        return out

    def lvg(self, ab, mask):
        """the escape-probability branch (3DPDR.F90:688-850)"""
        co_average = self.pct[0] == 100 and self.pct[1] == 100 and self.pct[2] == 100
        for p in np.where(mask)[0]:
            dens = ab[p, self.slots[4:]] * self.rho[p]      # H, H+, e-, He, H2
            for ci, tab in enumerate(self.tabs):
                a, b = self.off[ci], self.off[ci + 1]
                own = self.pop[p, a:b].copy()
                r = O.lvg_point(tab, self.T[p], self.Td[p], self.rho[p], self.phys.metallicity, self.phys.v_turb,
                                self.phys.g2d, dens, ab[p, self.slots[ci]], own, self.chains(p, ci),
                                pseudo_ray=self.pseudo_ray)
                s = r["solution"]
                if self.co_fix and (ci < 3 or co_average):
                    if self.levpop_iteration >= 120:
                        s = own
                    elif self.levpop_iteration >= 75:
                        s = 0.5 * (s + own)
                self.sol[p, a:b] = s
                self.cool[p, ci] = r["cooling"]
                for il in range(self.nlev[ci]):
                    for jl in range(il):
                        k = line_slot(ci, il, jl)
                        if k >= 0:
                            self.lines[p, k] = r["line"][il, jl]
                            self.tau_lines[p, k] = r["tau"][il, jl, self.tau_ray]
        self.total = ((self.cool[:, 0] + self.cool[:, 1]) + self.cool[:, 2]) + self.cool[:, 3]

    # -- temperature --------------------------------------------------------
    def update_temperature(self, p, fmean, fratio):
        """3DPDR.F90:936-1052 for one point"""
        Tmin, Tmax = self.Tmin, self.Tmax
        T = self.T[p]
        if self.first_time:
            dummy = T
            if fmean == 0:
                self.Tlow[p] = T
                self.Thigh[p] = T
            elif fmean > 0:
                self.Tlow[p] = T
                T = 1.3 * T
                self.prev[p] = "H"
            elif fmean < 0:
                self.Thigh[p] = T
                T = 0.7 * T
                if T < Tmin:
                    T = Tmin
                self.prev[p] = "C"
            self.Tprev[p] = Tmin if dummy < Tmin else dummy
            self.T[p] = T
            return
        dummy = T
        if fratio <= self.Fcrit:
            self.converged[p] = True
        if not self.binarychop[p]:
            if fmean > 0 and self.prev[p] == "H":
                self.Tlow[p] = T
                T = 1.3 * T
                self.Thigh[p] = T
                self.prev[p] = "H"
            if fmean < 0 and self.prev[p] == "C":
                self.Thigh[p] = T
                T = 0.7 * T
                self.Tlow[p] = T
                self.prev[p] = "C"
                if T < Tmin:
                    T = Tmin
                    self.Tlow[p] = Tmin
                    self.Thigh[p] = Tmin
            if fmean > 0 and self.prev[p] == "C":
                T = (self.Thigh[p] + self.Tlow[p]) / 2.0
                self.binarychop[p] = True
            if fmean < 0 and self.prev[p] == "H":
                T = (self.Thigh[p] + self.Tlow[p]) / 2.0
                self.binarychop[p] = True
        else:
            if fmean > 0:
                self.Tlow[p] = T
                T = (T + self.Thigh[p]) / 2.0
            if fmean < 0:
                self.Thigh[p] = T
                T = (T + self.Tlow[p]) / 2.0
        if self.temp_fix and abs(T - self.Tprev[p]) <= self.Tdiff and fratio > self.Fcrit:
            self.converged[p] = True
        self.Tprev[p] = dummy
        if dummy < Tmin and fmean < 0:
            self.converged[p] = True
        if dummy > Tmax and fmean > 0:
            self.converged[p] = True
        if self.converged[p]:
            if dummy < Tmin:
                self.Tprev[p] = Tmin
                T = Tmin
                if self.doleveltmin[p]:
                    self.converged[p] = True
                else:
                    self.converged[p] = False
                    self.level_converged[p] = False
                    dummy = Tmin
                self.doleveltmin[p] = True
            if dummy > Tmax:
                self.Tprev[p] = Tmax
                T = Tmax
        self.T[p] = T

    # -- one ITERATION ------------------------------------------------------
    def chemistry_due(self) -> bool:
        """whether the next iteration starts with the three chemistry
        iterations (3DPDR.F90:602)"""
        return not self.done and self.iteration + 1 > 1 and self.levpop_iteration + 1 == 1

    def iterate(self, ab, cols) -> dict:
        """One ITERATION given the abundances and columns (per ray: H2, HD,
        CO, C, S) as this iteration's chemistry left them."""
        ab = np.asarray(ab, dtype=np.float64)
        if self.done:
            return {"done": 1}
        if not self.started:
            self.lte(ab, np.ones(self.n, bool))
            self.started = True
        self.iteration += 1
        self.levpop_iteration += 1
        lvg = not (self.iteration > 1 and self.levpop_iteration == 1)
        if not lvg:
            self.lte(ab, ~(self.level_converged | self.converged))
        else:
            self.lvg(ab, ~(self.level_converged | self.converged))
        act = np.where(~self.converged)[0]
        if len(act):
            rate = self.orc.rates(self.T[act], self.Td[act], self.G[act], self.Av[act], cols[act], phys=self.phys)
            h = O.heating(self.name, self.net, self.rho[act], self.T[act], self.Td[act], self.uv[act], ab[act], rate,
                          phys=self.phys)
            for k, p in enumerate(act):
                self.heat[p] = h[k]
                fmean = h[k, 11] - self.total[p]
                fratio = 2.0 * abs(fmean) / abs(h[k, 11] + self.total[p])
                if self.level_conv:
                    self.update_temperature(p, fmean, fratio)
        if self.level_conv and self.first_time:
            self.first_time = False
        nconv = int(self.converged.sum())
        if self.level_conv:
            self.level_conv = False
        out = {"lvg": int(lvg), "thermal_points": len(act), "thermally_converged": nconv}
        if nconv == self.n:
            self.done = True
            out["done"] = 1
        else:
            cconv = np.ones((self.n, 4), bool)
            for p in range(self.n):
                for ci in range(4):
                    x = ab[p, self.slots[ci]]
                    for l in range(self.off[ci], self.off[ci + 1]):
                        s, o = self.sol[p, l], self.pop[p, l]
                        if s >= x * 1.0e-10:
                            relch = 0.0 if (s == 0.0 and o == 0.0) else 2.0 * abs((s - o) / (s + o))
                            if relch > 1.0e-2:
                                cconv[p, ci] = False
            self.level_converged = cconv.all(axis=1)
            if self.level_converged.all():
                self.levpop_iteration = 0
                self.level_conv = True
            nlc = int(self.level_converged.sum())
            self.pct = [int(100.0 * float(cconv[:, ci].sum()) / float(self.n)) for ci in range(4)]
            if int(100.0 * float(nlc) / float(self.n)) >= 100:
                self.level_converged[:] = False
            self.pop = self.sol.copy()
            out.update(done=0, level_converged=nlc, coolant_converged=[int(cconv[:, ci].sum()) for ci in range(4)])
        out.update(iteration=self.iteration, levpop_iteration=self.levpop_iteration, level_conv=int(self.level_conv),
                   first_time=int(self.first_time))
        return out

"""BOSampleStepper's self-consistent loop on the device
(src/BOSampleStepper.cpp:782-977) for a fixed geometry:

    for itscf in range(nitscf):
        update_density; charge mixing (only when nite > 0); update_vhxc
        repeat max(nite, 1) times, stopping early once the eigenvalue sum
        changes by less than delta_ratio * delta_ehart:
            energy(compute_hpsi) ; eigenvalue_sum = <c|H c> ; stepper update

Every step is enqueued through the C ABI without host synchronization; the
host reads back only the scalars the loop's decisions need (ehart and the
eigenvalue sum).  With ``graphs=True`` the first iteration runs eagerly
(it allocates the workspaces) and the two fixed launch sequences -- density,
mixing and potential; energy and stepper update -- are captured once and
replayed for every later iteration.
"""
from __future__ import annotations

import time

from ._lib import Context, DArray, EnergyFunctional, Graph, Mixer, WfStepper


class ScfLoop:
    def __init__(self, ef: EnergyFunctional, c: DArray, hpsi: DArray, wf_dyn="psda",
                 param=0.0, alpha=0.5, rc_kerker=10.0, ndim=3, rc1=3.0, prec="fp64",
                 graphs=True):
        self.ef = ef
        self.ctx: Context = ef.ctx
        self.c = c
        self.hpsi = hpsi
        self.prec = prec
        # one stepper per spin (each spin's bands form their own SlaterDet)
        sd = ef.sysdef
        if getattr(sd, "nspin", 1) == 2:
            u = int(sd.nst_up)
            self.sets = [(0, u), (u, sd.nst - u)]
            self.steppers = [WfStepper(ef, wf_dyn, param, nst=n) for _, n in self.sets if n > 0]
            self.sets = [x for x in self.sets if x[1] > 0]
        else:
            self.sets = [(0, sd.nst)]
            self.steppers = [WfStepper(ef, wf_dyn, param)]
        # both spins' densities are mixed as one vector
        self.mixer = Mixer(ef, alpha, rc_kerker, ndim, rc1)
        self.stepper = self.steppers[0]
        self.graphs = graphs
        self._g_pot = {}
        self.itscf = 0

    # the two launch sequences of one SCF iteration
    def _potential(self, mix: bool):
        self.ef.update_density(self.c, self.prec)
        if mix:
            self.mixer.update()
        self.ef.update_vhxc(False)
        if mix:
            for st in self.steppers:
                st.preprocess()

    def _step(self, first: bool):
        # the first step of an iteration sees the bands update_density read
        self.ef.energy_async(self.c, self.hpsi, forces=False, stress=False, prec=self.prec,
                             same_c=first)
        for st, (b0, _) in zip(self.steppers, self.sets):
            st.update(self.c, self.hpsi, band0=b0)

    def _run(self, key, fn, *args):
        if not self.graphs or self.itscf == 0:
            fn(*args)
            return
        if key not in self._g_pot:
            with Graph(self.ctx) as g:
                fn(*args)
            self._g_pot[key] = g
        self._g_pot[key].launch()

    def run(self, nitscf: int, nite: int, delta_ratio: float = 0.01):
        """Returns one dict per SCF iteration: etotal (of the last energy
        evaluated), ehart, the eigenvalue sums and the number of steps."""
        self.mixer.restart()
        for st in self.steppers:
            st.preprocess()
        self.itscf = 0
        history = []
        ehart = delta_ehart = 0.0
        for itscf in range(nitscf):
            self.itscf = itscf
            t0 = time.perf_counter()
            self._run(("pot", nite > 0), self._potential, nite > 0)
            eigs = []
            res = None
            for ite in range(max(nite, 1)):
                self._run(("step", ite == 0), self._step, ite == 0)
                res = self.ef.results()
                ws = [st.results() for st in self.steppers]
                w = dict(ws[0], eigenvalue_sum=sum(x["eigenvalue_sum"] for x in ws))
                if ite == 0:
                    if itscf == 0:
                        ehart, delta_ehart = res["ehart"], 0.0
                    elif itscf == 1:
                        ehart, delta_ehart = res["ehart"], abs(res["ehart"] - ehart)
                    else:
                        delta_ehart = min(delta_ehart, abs(res["ehart"] - ehart))
                        ehart = res["ehart"]
                eigs.append(w["eigenvalue_sum"])
                if ite > 0 and abs(eigs[-1] - eigs[-2]) < delta_ratio * delta_ehart:
                    break
            history.append(dict(etotal=res["etotal"], ehart=res["ehart"], eigenvalue_sum=eigs,
                                steps=len(eigs), energies=res,
                                seconds=time.perf_counter() - t0))
        return history

    def release(self):
        for g in self._g_pot.values():
            g.release()
        self._g_pot = {}
        for st in self.steppers:
            st.release()
        if self.mixer is not None:
            self.mixer.release()

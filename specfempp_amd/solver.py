"""The reference's solver-side interface, driven through the C ABI's seams.

These classes carry the names, call order, return values and error behaviour
of the SPECFEM++ interfaces they mirror, so that a workflow written against
the reference's solver runs unchanged on the MI355X engine, one seam call at
a time:

===============================  ===============================================
class                            reference
===============================  ===============================================
``Newmark``                      ``specfem::time_scheme::newmark<forward>`` and
                                 ``<combined>`` (include/timescheme/newmark.hpp,
                                 newmark.tpp:119-242; base class
                                 timescheme.hpp:86-215)
``DomainKernels``                ``kokkos_kernels::domain_kernels<wavefield>``
                                 (include/kokkos_kernels/domain_kernels.hpp:34-195)
``FrechetKernels``               ``kokkos_kernels::frechet_kernels``
                                 (include/kokkos_kernels/frechet_kernels.hpp:19-53)
``PeriodicTask``                 ``periodic_tasks::periodic_task``
                                 (include/periodic_tasks/periodic_task.hpp:15-66)
``TimeMarching``                 ``solver::time_marching<forward>::run`` and
                                 ``<combined>::run`` (include/solver/time_marching.tpp:9-247)
===============================  ===============================================

The engine's fused ``run()`` / ``adjoint_run()`` are the fast path of the same
loops; these classes are the drop-in path, one launch family per seam call.
"""
from __future__ import annotations

# medium tags in the order the reference's loops visit them
ACOUSTIC, ELASTIC_PSV, ELASTIC_SH, POROELASTIC, ELASTIC_PSV_T = (
    "acoustic", "elastic_psv", "elastic_sh", "poroelastic", "elastic_psv_t")


class Newmark:
    """Newmark time scheme of a forward or a combined (adjoint + backward)
    simulation.  ``apply_*_forward`` step the forward field (forward
    simulation) or the adjoint field (combined); ``apply_*_backward`` step the
    backward field with -dt and exist only in a combined simulation, as in the
    reference, whose forward scheme aborts there."""

    def __init__(self, nstep: int, nstep_between_samples: int, dt: float, simulation: str = "forward"):
        if simulation not in ("forward", "combined"):
            raise ValueError("simulation must be 'forward' or 'combined'")
        self.nstep, self.nstep_between_samples, self.dt = int(nstep), int(nstep_between_samples), float(dt)
        self.simulation = simulation
        self.seismogram_timestep = 0
        self.engine = None

    # timescheme.hpp:110-160
    def iterate_forward(self):
        return ((istep, self.dt) for istep in range(self.nstep))

    def iterate_backward(self):
        return ((istep, self.dt) for istep in range(self.nstep - 1, -1, -1))

    def get_max_timestep(self) -> int:
        return self.nstep

    def increment_seismogram_step(self):
        self.seismogram_timestep += 1

    def compute_seismogram(self, istep: int) -> bool:
        return istep % self.nstep_between_samples == 0

    def get_seismogram_step(self) -> int:
        return self.seismogram_timestep

    def get_max_seismogram_step(self) -> int:
        return self.nstep // self.nstep_between_samples

    def get_nstep_between_samples(self) -> int:
        return self.nstep_between_samples

    # the pure virtuals (timescheme.hpp:173-207)
    def link_assembly(self, engine):
        self.engine = engine

    def get_timestep(self) -> float:
        return self.dt

    def timescheme(self) -> str:
        return "Newmark"

    def print(self, out=None):
        text = (f"  Time Scheme:\n  ------------------------------\n  - Newmark\n"
                f"    simulation type = {self.simulation}\n    dt = {self.dt}\n    number of time steps = {self.nstep}\n")
        if out is not None:
            out.write(text)
        return text

    def _forward_field(self):
        return "forward" if self.simulation == "forward" else "adjoint"

    def _combined_only(self):
        if self.simulation != "combined":
            raise RuntimeError("backward phases exist only in a combined simulation")

    def apply_predictor_phase_forward(self, tag: str) -> int:
        return self.engine.predictor_phase(tag, self._forward_field(), "forward")

    def apply_corrector_phase_forward(self, tag: str) -> int:
        return self.engine.corrector_phase(tag, self._forward_field(), "forward")

    def apply_predictor_phase_backward(self, tag: str) -> int:
        self._combined_only()
        return self.engine.predictor_phase(tag, "backward", "backward")

    def apply_corrector_phase_backward(self, tag: str) -> int:
        self._combined_only()
        return self.engine.corrector_phase(tag, "backward", "backward")


class DomainKernels:
    """domain_kernels of one wavefield: initialize (mass matrix, zero field),
    update_wavefields<medium>(istep) (returns the medium's elements) and
    compute_seismograms(isig)."""

    def __init__(self, wavefield: str, engine):
        self.wavefield, self.engine = wavefield, engine

    def initialize(self, dt: float):
        self.engine.initialize_wavefield(self.wavefield)

    def update_wavefields(self, medium: str, istep: int) -> int:
        return self.engine.update_wavefields_medium(medium, self.wavefield, istep)

    def compute_seismograms(self, isig: int):
        self.engine.compute_seismograms_wavefield(self.wavefield, isig)


class FrechetKernels:
    def __init__(self, engine):
        self.engine = engine

    def compute_derivatives(self, dt: float):
        self.engine.compute_derivatives(dt)


class PeriodicTask:
    """Base of tasks run by the time loop every ``period`` steps (a wavefield
    writer, a plotter, a signal check)."""

    def __init__(self, period: int = 1):
        self.period = int(period)

    def should_run(self, istep: int) -> bool:
        return istep % self.period == 0 if self.period > 0 else False

    def initialize(self, engine):
        pass

    def run(self, engine, istep: int):
        pass

    def finalize(self, engine):
        pass


class TimeMarching:
    """time_marching<simulation>::run over an engine and a ``Newmark`` scheme
    linked to it, with the reference's accounting: every step must update
    2 (forward) or 4 (combined) times the degrees of freedom and 1 or 2 times
    the elements, or ``RuntimeError`` is raised (time_marching.tpp:90-108,
    215-235)."""

    def __init__(self, engine, time_scheme: Newmark, tasks=(), progress=None):
        self.engine, self.time_scheme, self.tasks = engine, time_scheme, list(tasks)
        self.progress = progress
        # assembly.get_total_degrees_of_freedom(): every medium's field points
        # times its components (the acoustic potential has one)
        self.total_dofs = (getattr(engine, "field_nglob", engine.nglob) * engine.ncomp
                           + getattr(engine, "nglob_acoustic", 0))
        self.total_elements = engine.nspec

    def _check(self, dofs, elements, factor):
        if dofs != 2 * factor * self.total_dofs:
            raise RuntimeError(f"The time loop has not updated all the degrees of freedom. Only {dofs} out of "
                               f"{2 * factor * self.total_dofs} degrees of freedom have been updated.")
        if elements != factor * self.total_elements:
            raise RuntimeError(f"The time loop has not updated all the elements. Only {elements} out of "
                               f"{factor * self.total_elements} elements have been updated.")

    def run(self):
        return self._run_combined() if self.time_scheme.simulation == "combined" else self._run_forward()

    def _run_forward(self):
        ts, kernels = self.time_scheme, DomainKernels("forward", self.engine)
        kernels.initialize(ts.get_timestep())
        nstep = ts.get_max_timestep()
        for task in self.tasks:
            task.initialize(self.engine)
        for istep, _dt in ts.iterate_forward():
            dofs = sum(ts.apply_predictor_phase_forward(m)
                       for m in (ACOUSTIC, ELASTIC_PSV, ELASTIC_SH, POROELASTIC, ELASTIC_PSV_T))
            elements = kernels.update_wavefields(ACOUSTIC, istep)
            dofs += ts.apply_corrector_phase_forward(ACOUSTIC)
            for m in (ELASTIC_PSV, ELASTIC_SH, ELASTIC_PSV_T):
                elements += kernels.update_wavefields(m, istep)
            for m in (ELASTIC_PSV, ELASTIC_SH, ELASTIC_PSV_T):
                dofs += ts.apply_corrector_phase_forward(m)
            elements += kernels.update_wavefields(POROELASTIC, istep)
            dofs += ts.apply_corrector_phase_forward(POROELASTIC)
            if ts.compute_seismogram(istep):
                kernels.compute_seismograms(ts.get_seismogram_step())
                ts.increment_seismogram_step()
            for task in self.tasks:
                if task.should_run(istep + 1):
                    task.run(self.engine, istep + 1)
            if self.progress and istep % 10 == 0:
                self.progress(f"Progress : executed {istep} steps of {nstep} steps")
            self._check(dofs, elements, 1)
        for task in self.tasks:
            if not task.should_run(nstep) and task.should_run(-1):
                task.run(self.engine, nstep)
        for task in self.tasks:
            task.finalize(self.engine)

    def _run_combined(self):
        """The buffer wavefield must hold the forward run's final state and the
        engine the forward run's Stacey record (a forward run's wavefield
        writer, or ``forward_record`` followed by ``copy_wavefield("buffer",
        "forward")``, or the checkpoint reader)."""
        ts = self.time_scheme
        adjoint, backward = DomainKernels("adjoint", self.engine), DomainKernels("backward", self.engine)
        frechet = FrechetKernels(self.engine)
        adjoint.initialize(ts.get_timestep())
        backward.initialize(ts.get_timestep())
        nstep = ts.get_max_timestep()
        for task in self.tasks:
            task.initialize(self.engine)
        for task in self.tasks:
            if not task.should_run(nstep) and task.should_run(-1):
                task.run(self.engine, nstep)
        for istep, dt in ts.iterate_backward():
            for task in self.tasks:
                if task.should_run(istep + 1):
                    task.run(self.engine, istep + 1)
            # adjoint step
            dofs = sum(ts.apply_predictor_phase_forward(m) for m in (ACOUSTIC, ELASTIC_PSV, ELASTIC_SH, POROELASTIC))
            elements = adjoint.update_wavefields(ACOUSTIC, istep)
            dofs += ts.apply_corrector_phase_forward(ACOUSTIC)
            elements += adjoint.update_wavefields(ELASTIC_PSV, istep)
            elements += adjoint.update_wavefields(ELASTIC_SH, istep)
            dofs += ts.apply_corrector_phase_forward(ELASTIC_PSV)
            dofs += ts.apply_corrector_phase_forward(ELASTIC_SH)
            elements += adjoint.update_wavefields(POROELASTIC, istep)
            dofs += ts.apply_corrector_phase_forward(POROELASTIC)
            # backward step
            dofs += sum(ts.apply_predictor_phase_backward(m) for m in (ELASTIC_PSV, ACOUSTIC, POROELASTIC))
            elements += backward.update_wavefields(ELASTIC_PSV, istep)
            elements += backward.update_wavefields(ELASTIC_SH, istep)
            dofs += ts.apply_corrector_phase_backward(ELASTIC_PSV)
            dofs += ts.apply_corrector_phase_backward(ELASTIC_SH)
            elements += backward.update_wavefields(ACOUSTIC, istep)
            dofs += ts.apply_corrector_phase_backward(ACOUSTIC)
            elements += backward.update_wavefields(POROELASTIC, istep)
            dofs += ts.apply_corrector_phase_backward(POROELASTIC)
            # the backward field joins the forward run's final state after its
            # first step (time_marching.tpp:201-209)
            if istep == nstep - 1:
                self.engine.copy_wavefield("backward", "buffer")
            frechet.compute_derivatives(dt)
            if ts.compute_seismogram(istep):
                backward.compute_seismograms(ts.get_seismogram_step())
                ts.increment_seismogram_step()
            if self.progress and istep % 10 == 0:
                self.progress(f"Progress : executed {istep} steps of {nstep} steps")
            self._check(dofs, elements, 2)
        for task in self.tasks:
            task.finalize(self.engine)

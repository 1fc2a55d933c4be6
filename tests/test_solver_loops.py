"""Host logic of the seam-driven loops (specfempp_amd.solver) on a stand-in
engine that records the calls: the call order of time_marching.tpp:9-247,
the accounting errors, and the forward scheme's refusal of backward phases
(the reference's newmark<forward> has no backward phase)."""
import pytest

from specfempp_amd.solver import Newmark, PeriodicTask, TimeMarching


class FakeEngine:
    """Elastic P-SV mesh of 10 elements and 50 points: 100 DOFs."""

    def __init__(self, lose_dofs=False):
        self.nglob, self.ncomp, self.nspec, self.calls = 50, 2, 10, []
        self.lose_dofs = lose_dofs

    def _dofs(self, medium):
        return 0 if medium != "elastic_psv" else (99 if self.lose_dofs else 100)

    def initialize_wavefield(self, w):
        self.calls.append(("init", w))

    def predictor_phase(self, m, w, d):
        self.calls.append(("pred", m, w, d))
        return self._dofs(m)

    def corrector_phase(self, m, w, d):
        self.calls.append(("corr", m, w, d))
        return self._dofs(m)

    def update_wavefields_medium(self, m, w, istep):
        self.calls.append(("update", m, w, istep))
        return 10 if m == "elastic_psv" else 0

    def compute_seismograms_wavefield(self, w, isig):
        self.calls.append(("seismo", w, isig))

    def copy_wavefield(self, dst, src):
        self.calls.append(("copy", dst, src))

    def compute_derivatives(self, dt):
        self.calls.append(("frechet", dt))


def test_forward_loop_call_order_and_samples():
    e = FakeEngine()
    ts = Newmark(4, 2, 1e-3)
    ts.link_assembly(e)
    TimeMarching(e, ts).run()
    elastic = [c for c in e.calls if c[0] != "init" and (len(c) < 2 or c[1] in ("elastic_psv", "forward"))]
    assert elastic[:4] == [("pred", "elastic_psv", "forward", "forward"), ("update", "elastic_psv", "forward", 0),
                           ("corr", "elastic_psv", "forward", "forward"), ("seismo", "forward", 0)]
    assert [c for c in e.calls if c[0] == "seismo"] == [("seismo", "forward", 0), ("seismo", "forward", 1)]
    assert ts.get_seismogram_step() == 2


def test_combined_loop_call_order():
    e = FakeEngine()
    ts = Newmark(3, 1, 1e-3, "combined")
    ts.link_assembly(e)
    TimeMarching(e, ts).run()
    steps = [c[3] for c in e.calls if c[0] == "update" and c[1] == "elastic_psv" and c[2] == "adjoint"]
    assert steps == [2, 1, 0]
    first = e.calls.index(("update", "elastic_psv", "backward", 2))
    assert e.calls[first + 1:first + 4] == [("update", "elastic_sh", "backward", 2),
                                            ("corr", "elastic_psv", "backward", "backward"),
                                            ("corr", "elastic_sh", "backward", "backward")]
    assert [c for c in e.calls if c[0] == "copy"] == [("copy", "backward", "buffer")]
    copy_at = e.calls.index(("copy", "backward", "buffer"))
    assert e.calls[copy_at + 1] == ("frechet", 1e-3)
    assert e.calls.count(("frechet", 1e-3)) == 3


def test_accounting_errors_and_forward_scheme_has_no_backward_phase():
    e = FakeEngine(lose_dofs=True)
    ts = Newmark(2, 1, 1e-3)
    ts.link_assembly(e)
    with pytest.raises(RuntimeError, match="degrees of freedom"):
        TimeMarching(e, ts).run()
    with pytest.raises(RuntimeError, match="combined"):
        Newmark(2, 1, 1e-3).apply_predictor_phase_backward("elastic_psv")


def test_periodic_tasks_run_at_their_steps():
    class Recorder(PeriodicTask):
        def __init__(self):
            super().__init__(period=2)
            self.seen = []

        def run(self, engine, istep):
            self.seen.append(istep)

    e, task = FakeEngine(), Recorder()
    ts = Newmark(5, 1, 1e-3)
    ts.link_assembly(e)
    TimeMarching(e, ts, tasks=[task]).run()
    assert task.seen == [2, 4]

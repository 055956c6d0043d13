"""Callbacks (pylops/optimization/callback.py): hooks around setup, step and
run of a solver."""

__all__ = ["Callbacks"]


class Callbacks:
    """Subclass and override any of the hooks; each receives the solver and
    the current model vector (``x0`` for setup)."""

    def on_setup_begin(self, solver, x0):
        pass

    def on_setup_end(self, solver, x):
        pass

    def on_step_begin(self, solver, x):
        pass

    def on_step_end(self, solver, x):
        pass

    def on_run_begin(self, solver, x):
        pass

    def on_run_end(self, solver, x):
        pass

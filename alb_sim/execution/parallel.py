from pyalbsim_amd.execution.parallel import run_parallel  # noqa: F401

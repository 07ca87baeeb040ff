from pyalbsim_amd.execution.linear import run_linear, simulate  # noqa: F401

from pyalbsim_amd.config.simulation import *  # noqa: F401,F403

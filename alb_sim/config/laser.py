from pyalbsim_amd.config.laser import *  # noqa: F401,F403

from pyalbsim_amd.config.water import *  # noqa: F401,F403

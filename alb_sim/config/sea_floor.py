from pyalbsim_amd.config.sea_floor import *  # noqa: F401,F403
